qreg q[3];
gate two a, b { cz a, b; }
two q[0], q[1], q[2];
