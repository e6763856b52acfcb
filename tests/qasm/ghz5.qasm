// 5-qubit GHZ state written with a user gate definition.
OPENQASM 2.0;
include "qelib1.inc";
gate entangle a, b { cx a, b; }
qreg q[5];
creg c[5];
h q[0];
entangle q[0], q[1];
entangle q[1], q[2];
entangle q[2], q[3];
entangle q[3], q[4];
measure q -> c;
