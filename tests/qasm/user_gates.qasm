// Gate definitions: parameters, nesting, use before definition, a repeated
// definition (the first one is kept), and the qelib1 aliases.
OPENQASM 2.0;
include "qelib1.inc";
qreg data[3];
qreg anc[2];
creg out[3];
creg flag[1];

layer(pi/3, -0.25) data[0], data[2], anc[1];

gate rot(alpha) t { rz(alpha) t; ry(alpha / 2) t; }
gate layer(a, b) x, y, z {
  rot(a) x;
  cx x, y;
  rot(b * 2) z;
  cp(a - b) z, x;
}
gate rot(alpha) t { rx(alpha) t; }
gate swapish p, q { cx p, q; cx q, p; cx p, q; }

swapish anc[0], data[1];
u2(0, pi) data[1];
u1(pi/8) anc[0];
p(-pi/16) anc[1];
phase(3*pi/4) data[0];
id data[2];
U(0.1, 0.2, 0.3) data[2];
u3(1e-3, 2.5e1, -.5) anc[0];
barrier data, anc[1];
layer(0.5, 0.5) anc[1], anc[0], data[1];
measure data[0] -> out[2];
measure anc[1] -> flag[0];
