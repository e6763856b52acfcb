"""OpenQASM 2 front end: this backend's from_qasm / to_qasm against ket's own
(bindings/qasm/qasm.cpp, compiled from the reference sources into
oracle/_ref/qasm_tool, and its outputs committed under tests/golden/).

The same tests/cpp/qasm_tool.cpp is built against both libraries. Parsing and
writing need no GPU, so the byte comparisons run on the CPU; the behaviour
cases that simulate (ket::run) run on the GPU.
"""
import glob
import hashlib
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MINE = os.path.join(ROOT, "ket_amd", "qasm_tool")
REF = os.path.join(ROOT, "oracle", "_ref", "qasm_tool")
GOLDEN = os.path.join(ROOT, "tests", "golden")
FILES = sorted(glob.glob(os.path.join(ROOT, "tests", "qasm", "*.qasm")))
# OpenQASM 2 forms ket's parser rejects and this one accepts: register
# broadcast (`h q;`, `measure q -> c;`) and `^` in angles.
SUPERSET = {"ghz5.qasm", "grover3.qasm", "qft_basis.qasm"}
ref_built = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")


def tool(exe, *args, check=True):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120)
    if check:
        assert r.returncode == 0, r.stdout + r.stderr
    return r


def sections(text):
    out, name = {}, None
    for line in text.splitlines(keepends=True):
        if line.startswith("== "):
            name = line[3:].strip()
            out[name] = ""
        else:
            out[name] += line
    return out


def check_dump(mine, ref):
    m, r = sections(mine), sections(ref)
    assert sorted(m) == sorted(r)
    for name in m:
        if name in SUPERSET:
            assert r[name] == "error: runtime_error\n", name
            assert "fixed_point=1" in m[name], name
        else:
            assert m[name] == r[name], name


def test_dump_equals_reference_fixture():
    with open(os.path.join(GOLDEN, "qasm_dump_ref.txt")) as f:
        want = f.read()
    check_dump(tool(MINE, "dump", *FILES).stdout, want)
    # Inputs that both parsers take re-serialize as fixed points; the error
    # files fail in both.
    got = sections(want)
    assert sum("fixed_point=1" in v for v in got.values()) == 2
    assert sum(v == "error: runtime_error\n" for v in got.values()) == 7


@ref_built
def test_dump_equals_live_reference():
    ref = tool(REF, "dump", *FILES).stdout
    with open(os.path.join(GOLDEN, "qasm_dump_ref.txt")) as f:
        assert f.read() == ref, "fixture is stale: rerun tests/golden/make_qasm_golden.py"
    check_dump(tool(MINE, "dump", *FILES).stdout, ref)


def test_generated_circuits_equal_reference_fixture():
    with open(os.path.join(GOLDEN, "qasm_gen_ref.json")) as f:
        g = json.load(f)
    for seed, digest in g["sha256"].items():
        out = tool(MINE, "gen", seed, g["per_seed"]).stdout
        assert hashlib.sha256(out.encode()).hexdigest() == digest, seed
        # Every generated circuit's text is a fixed point of parse + write.
        assert out.count("fixed_point=1") == g["per_seed"], seed
        assert "fixed_point=0" not in out and "reparse:" not in out


@ref_built
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_generated_circuits_equal_live_reference(seed):
    assert tool(MINE, "gen", seed, 60).stdout == tool(REF, "gen", seed, 60).stdout


@ref_built
def test_cases_pass_on_reference():
    """The restated cases hold for ket itself (this checks the cases)."""
    r = tool(REF, "cases", check=False)
    assert r.returncode == 0, r.stdout
    assert sum(line.startswith("ok ") for line in r.stdout.splitlines()) == 24


def test_cases_without_simulation():
    r = tool(MINE, "cases", "--no-run", check=False)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert sum(line.startswith("ok ") for line in r.stdout.splitlines()) == 15


@pytest.mark.gpu
def test_cases_with_simulation():
    r = tool(MINE, "cases", check=False)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert sum(line.startswith("ok ") for line in r.stdout.splitlines()) == 24


@ref_built
@pytest.mark.skipif(not os.path.isdir("/root/reference/examples"), reason="reference tree absent")
def test_reference_example_programs():
    ex = sorted(glob.glob("/root/reference/examples/*.qasm"))
    assert ex
    assert tool(MINE, "dump", *ex).stdout == tool(REF, "dump", *ex).stdout
