"""CPU-side checks of the drop-in boundary: libhpcg_amd.so builds, loads and
exports every entry point declared in include/hpcg_amd.h; the product's
flop/byte model agrees with the oracle's restatement of ReportResults.cpp.
No compute call is made here (there is no GPU in this suite)."""
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "hpcg_amd.h")
LIB = os.path.join(ROOT, "hpcg_amd", "libhpcg_amd.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(hpcgx_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first: python -m hpcg_amd.build"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (hpcgx_\w+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    from hpcg_amd import exported_symbols
    assert sorted(exported_symbols()) == declared()


def test_library_loads_and_reports_version():
    import hpcg_amd
    L = hpcg_amd.lib()
    assert L.hpcgx_abi_version() == 1
    for s in declared():
        assert hasattr(L, s)


def test_model_matches_oracle_model():
    from hpcg_amd import model
    from oracle import pyoracle as O
    for dims in [(16, 16, 16), (256, 256, 256), (512, 512, 640)]:
        lv = model.grid_levels(*dims)
        nrow = [l["n"] for l in lv]
        nnz = [l["nnz"] for l in lv]
        a = model.hpcg_model(nrow, nnz, 3, 50, 52)
        b = O.hpcg_model(nrow, nnz, 3, 50, 52)
        for k in a:
            assert a[k] == pytest.approx(b[k], rel=1e-15), (dims, k)


def test_grid_levels_match_oracle_generator():
    from hpcg_amd import model
    from oracle import pyoracle as O
    import numpy as np
    for dims in [(16, 16, 16), (8, 16, 24), (32, 24, 48)]:
        h = O.Hierarchy(*dims)
        lv = model.grid_levels(*dims)
        for l in range(4):
            assert lv[l]["n"] == h.nrow(l) and lv[l]["nnz"] == h.nnz(l)
            if l < 3:
                m = h.level(l)
                f2c = m["f2c"]
                assert lv[l]["m_inj"] == int(m["nnz_row"][f2c].astype(np.int64).sum())


def test_canonical_bytes_closed_forms():
    """Per-iteration compulsory bytes quoted for the 256^3 and 512x512x640
    FP64 configurations (3.500e10 B and 3.514e11 B)."""
    from hpcg_amd import model
    b = model.canonical_bytes(model.grid_levels(256, 256, 256))
    assert b["spmv_L0"] == 12 * 449455096 + 16 * 16777216
    assert b["iteration"] == pytest.approx(3.500e10, rel=1e-3)
    b = model.canonical_bytes(model.grid_levels(512, 512, 640))
    assert b["iteration"] == pytest.approx(3.514e11, rel=1e-3)


def test_tune_keys_match_header():
    """Every key hpcgx_tune accepts is documented in the header's key list,
    and every documented key is accepted (the key table is the source)."""
    doc = set(re.findall(r'^\s*\*\s+"([a-z0-9_]+)"', open(HEADER).read(), re.M))
    src = open(os.path.join(ROOT, "hpcg_amd", "csrc", "capi.hip")).read()
    accepted = set(re.findall(r'k == "([a-z0-9_]+)"', src))
    assert accepted and doc == accepted, sorted(doc ^ accepted)
