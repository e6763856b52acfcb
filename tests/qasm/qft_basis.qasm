// QFT of |0110> on two registers, with parameterised cu1 angles written as
// pi expressions; every outcome has probability 1/16.
OPENQASM 2.0;
include "qelib1.inc";
qreg a[2];
qreg b[2];
x a[1];
x b[0];
h a[0];
cu1(pi/2) a[1], a[0];
cu1(pi/4) b[0], a[0];
cu1(pi/2^3) b[1], a[0];
h a[1];
cu1(pi/2) b[0], a[1];
cu1(0.25*pi) b[1], a[1];
h b[0];
cu1(2*pi/4) b[1], b[0];
h b[1];
swap a[0], b[1];
swap a[1], b[0];
u3(0, 0, 0) a[0];
u2(0, pi) b[1];
u2(0, pi) b[1];
