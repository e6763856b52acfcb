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


def test_tuning_struct_matches_header():
    text = open(os.path.join(ROOT, "include", "ket_amd.h")).read()
    body = re.search(r"typedef struct ket_tuning \{(.*?)\} ket_tuning;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"int32_t\s+(\w+);", body)
    assert tuple(fields) == N.TUNING_FIELDS
    assert ctypes.sizeof(N.Tuning) == 4 * len(fields)


def test_tuning_round_trip_and_validation():
    """ket_set_tuning_for / ket_get_tuning_for keep each precision's fields
    and reject out-of-range values (no device call)."""
    import pytest
    saved = {p: N.get_tuning(p) for p in (N.C128, N.C64)}
    try:
        N.set_tuning(N.C64, block_k=3, pair_exchange=16)
        assert N.get_tuning(N.C64)["block_k"] == 3
        assert N.get_tuning(N.C64)["pair_exchange"] == 16
        assert N.get_tuning(N.C128) == saved[N.C128]
        for bad in ({"block_k": 1}, {"block_k": 7}, {"nontemporal": 4}, {"items_per_thread": 3},
                    {"block_stage": 65}, {"pair_exchange": -1}):
            with pytest.raises(ValueError):
                N.set_tuning(N.C128, **bad)
        assert N.get_tuning(N.C128) == saved[N.C128]
    finally:
        for p, t in saved.items():
            N.set_tuning(p, **t)


def test_default_tuning_is_the_recorded_autotuner_pick():
    """The library's built-in launch shapes are the picks the autotuner
    recorded on MI355X (ket_amd/tuning.json), field by field."""
    import json
    table = json.load(open(os.path.join(ROOT, "ket_amd", "tuning.json")))
    for p, key, width in ((N.C128, "c128", 4), (N.C64, "c64", 5)):
        got = N.get_tuning(p)
        for section in ("", "_blocks", "_width", "_runs"):
            for field, value in table[key + section]["best"].items():
                if field == "block_k":  # 0 selects the built-in width
                    assert value == width and got["block_k"] in (0, width), (key, field)
                else:
                    assert got[field] == value, (key, section, field, got[field], value)
