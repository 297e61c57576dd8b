import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


def load_golden(name):
    """Load tests/golden/ref_<name>.npz; per-row column offsets become
    absolute column indices again (`L<l>_cols`)."""
    d = dict(np.load(os.path.join(GOLDEN, "ref_%s.npz" % name)))
    for k in list(d):
        if k.endswith("_coldelta"):
            rows = np.repeat(np.arange(d[k].size // 27, dtype=np.int64), 27)
            d[k[:-9] + "_cols"] = (d.pop(k).astype(np.int64) + rows).astype(np.int32)
    return d


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
