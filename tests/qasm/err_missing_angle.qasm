qreg q[1];
rz q[0];
