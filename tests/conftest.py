import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def load_golden(name):
    """Load tests/golden/ref_<name>.npz; per-row column offsets become
    absolute column indices again (`L<l>_cols`)."""
    d = dict(np.load(os.path.join(GOLDEN, "ref_%s.npz" % name)))
    for k in list(d):
        if k.endswith("_coldelta"):
            rows = np.repeat(np.arange(d[k].size // 27, dtype=np.int64), 27)
            d[k[:-9] + "_cols"] = (d.pop(k).astype(np.int64) + rows).astype(np.int32)
    return d


def splitmix_uniform(seed, n):
    """splitmix64 stream mapped to uniform [-1, 1) (53-bit mantissa), the
    same stream oracle/ref_harness.cpp draws; vectorised with wrapping
    uint64 arithmetic."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * np.arange(1, n + 1, dtype=np.uint64)
        z = (s ^ (s >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def golden_level(g, l):
    return dict(nnz_row=g["L%d_nnz_row" % l], cols=g["L%d_cols" % l], vals=g["L%d_vals" % l],
                diag_slot=g["L%d_diag_slot" % l])


@pytest.fixture(scope="session")
def g16():
    return load_golden("16x16x16")


@pytest.fixture(scope="session")
def g81624():
    return load_golden("8x16x24")


@pytest.fixture(scope="session")
def g32():
    return load_golden("32x32x32")
