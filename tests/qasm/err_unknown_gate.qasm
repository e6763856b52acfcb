qreg q[2];
rzz(0.5) q[0], q[1];
