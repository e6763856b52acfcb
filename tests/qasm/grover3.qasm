// Grover search on 3 qubits for |101> (two iterations): p(|101>) = 0.9453125.
OPENQASM 2.0;
include "qelib1.inc";
gate ccz a, b, c { h c; ccx a, b, c; h c; }
gate oracle a, b, c { x b; ccz a, b, c; x b; }
gate diffuse a, b, c { h a; h b; h c; x a; x b; x c; ccz a, b, c; x a; x b; x c; h a; h b; h c; }
qreg q[3];
creg m[3];
h q;
oracle q[0], q[1], q[2];
diffuse q[0], q[1], q[2];
barrier q;
oracle q[0], q[1], q[2];
diffuse q[0], q[1], q[2];
measure q -> m;
