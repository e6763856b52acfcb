"""Writes the OpenQASM fixtures under tests/golden/ from the compiled reference.

Run where the reference tree is present and oracle/_ref is built
(`make -C oracle ref`). The texts come out of the reference's own
from_qasm / to_qasm (bindings/qasm/qasm.cpp) through tests/cpp/qasm_tool.cpp;
nothing is computed by this repository's code.

    python tests/golden/make_qasm_golden.py

qasm_dump_ref.txt   `qasm_tool dump` over tests/qasm/*.qasm
qasm_gen_ref.json   sha256 of `qasm_tool gen <seed> 40` for seeds 1..8
"""
from __future__ import annotations

import glob
import hashlib
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
TOOL = os.path.join(ROOT, "oracle", "_ref", "qasm_tool")
SEEDS = list(range(1, 9))
PER_SEED = 40


def qasm_files():
    return sorted(glob.glob(os.path.join(ROOT, "tests", "qasm", "*.qasm")))


def main():
    out = subprocess.run([TOOL, "dump"] + qasm_files(), check=True, capture_output=True,
                         text=True).stdout
    with open(os.path.join(HERE, "qasm_dump_ref.txt"), "w") as f:
        f.write(out)
    gen = {}
    for s in SEEDS:
        text = subprocess.run([TOOL, "gen", str(s), str(PER_SEED)], check=True,
                              capture_output=True).stdout
        gen[str(s)] = hashlib.sha256(text).hexdigest()
    with open(os.path.join(HERE, "qasm_gen_ref.json"), "w") as f:
        json.dump({"per_seed": PER_SEED, "sha256": gen}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
