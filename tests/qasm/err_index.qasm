qreg q[2];
h q[2];
