"""The C ABI library loads and exports every function include/ket_amd.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from ket_amd import native as N
from ket_amd.gates import GATE_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ket_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    names = declared("ket_amd.h")
    assert len(names) >= 20
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    names = set(declared("ket_amd.h"))
    N.lib()  # sets argtypes for every bound entry point
    src = open(os.path.join(ROOT, "ket_amd", "native.py")).read()
    unbound = [n for n in names if f'"{n}"' not in src]
    assert not unbound, unbound


def test_gate_record_layout_matches_header():
    text = open(os.path.join(ROOT, "include", "ket_amd.h")).read()
    assert "int32_t qubits[3];" in text and "double params[3];" in text
    assert GATE_DTYPE.itemsize == 48
    assert GATE_DTYPE.fields["p"][1] == 24
