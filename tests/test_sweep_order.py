"""The colour visiting order of the multicolour sweeps, pinned on the oracle
(CPU only).

On the 27-point grid the greedy first-fit colour of a point is its parity
class (x&1) + 2(y&1) + 4(z&1), and colour 0 (x, y, z all even) holds the
points the restriction injects: f2c maps coarse point (i, j, k) to fine
point (2i, 2j, 2k) (GenerateCoarseProblem.cpp:77-79) and the coarse
residual is rf - A xf there (ComputeRestriction_ref.cpp:51).  A symmetric
sweep in colour id order ends its backward half on colour 0, which leaves
the residual of those rows at zero up to rounding, so the coarse grid gets
almost nothing.  The default order (oracle SWEEP_ORDERS[0], the device's
HPCGX_SWEEP_ORDER(0)) visits colour 0 last forward and first backward.
"""
import numpy as np
import pytest

from oracle import pyoracle as O


def test_injected_rows_are_colour_zero():
    h = O.Hierarchy(16, 16, 16)
    for l in range(3):
        c, _ = O.colors(h.level(l))
        assert np.all(c[h.level(l)["f2c"]] == 0), l


@pytest.mark.parametrize("dims", [(16, 16, 16), (8, 12, 16)])
def test_id_order_leaves_zero_residual_where_restriction_injects(dims):
    """After one symmetric sweep the residual at the injected rows is zero up
    to rounding in id order, and of the size of the other rows' in the
    default order."""
    m = O.generate(*dims)
    n = int(np.prod(dims))
    c, _ = O.colors(m)
    inj = c == 0
    rng = np.random.default_rng(3)
    r, x0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    for sweep, zero in ((1, True), (0, False)):
        x = O.symgs(m, r, x0, O.color_order(c, sweep))
        res = np.abs(r - O.spmv(m, x))
        scale = np.abs(res[~inj]).max()
        if zero:
            assert np.abs(res[inj]).max() <= 1e-13 * (1 + np.abs(r).max()), sweep
        else:
            assert np.abs(res[inj]).max() > 1e-3 * scale, sweep


def test_color_order_ranks():
    c = np.array([0, 3, 7, 1, 0, 3, 2, 5, 6, 4, 9])
    o = O.color_order(c)
    assert np.array_equal(c[o], [3, 3, 5, 6, 7, 1, 2, 4, 0, 0, 9])
    assert np.array_equal(o[:2], [1, 5]) and np.array_equal(o[-3:-1], [0, 4])  # stable
    assert np.array_equal(c[O.color_order(c, 1)], np.sort(c))


def test_default_order_needs_fewer_pcg_iterations():
    """32x48x64: to the residual 50 lexicographic iterations reach
    (7.3e-15 relative) the id order needs 62 PCG iterations and the default
    53; to 1e-9, 42 and 36."""
    dims = (32, 48, 64)
    h = O.Hierarchy(*dims)
    ref = h.cg(max_iter=50, tol=0.0)
    tol = ref["normr"] / ref["normr0"]
    got = {}
    for sweep in (1, 0):
        h.use_color_order(sweep)
        got[sweep] = (h.cg(max_iter=300, tol=tol)["niters"], h.cg(max_iter=300, tol=1e-9)["niters"])
    assert got == {1: (62, 42), 0: (53, 36)}
