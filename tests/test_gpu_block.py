"""Block PCG over s right-hand sides (hpcgx_*_block) against the oracle.

Every right-hand side gets the single-vector arithmetic, so per column:
  * FP64 SpMV and the FP32 V-cycle (both orderings) are bit-exact against
    the oracle and against the single-vector GPU kernels;
  * the CG trajectory matches the oracle's mixed-precision CG on that column
    within 1e-9 relative before the first FP64 residual recompute (end of
    iteration 10) and 1e-7 from it on (the dots' reassociation, amplified by
    the recompute's cancellation, as in test_gpu_mixed.py).
Right-hand sides: splitmix64 uniform(-1, 1) from seed 42, one stream read
row-major into [n][s].
"""
import sys

import numpy as np
import pytest

from conftest import splitmix_uniform
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hp():
    import torch
    assert torch.cuda.is_available()
    import hpcg_amd
    return hpcg_amd


def rhs(n, s, seed=42):
    return splitmix_uniform(seed, n * s).reshape(n, s)


@pytest.mark.parametrize("dims,s", [((16, 16, 16), 8), ((32, 48, 64), 4), ((32, 48, 64), 16)])
def test_spmv_block_bitexact(hp, dims, s):
    say("spmv block", dims, s)
    P = hp.Problem.generate(*dims, nlevels=4)
    h = O.Hierarchy(*dims)
    for l in (0, 1):
        n = h.nrow(l)
        Xn = rhs(n, s, 7 + l)
        Y = P.to_device_order_multi(Xn, l) * 0
        P.spmv_block(P.to_device_order_multi(Xn, l), Y, level=l)
        got = P.to_numpy_multi(Y, l)
        for k in range(s):
            assert np.array_equal(got[:, k], O.spmv(h.level(l), Xn[:, k])), (l, k)
    del P


@pytest.mark.parametrize("ordering", ["multicolor", "lexicographic"])
@pytest.mark.parametrize("dims,s", [((16, 16, 16), 8), ((32, 48, 64), 8)])
def test_mg32_block_bitexact(hp, dims, s, ordering):
    import torch
    say("mg32 block", dims, s, ordering)
    P = hp.Problem.generate(*dims, nlevels=4)
    h = O.Hierarchy(*dims)
    if ordering == "multicolor":
        h.use_color_order()
    n = h.nrow(0)
    Rn = rhs(n, s, 11).astype(np.float32)
    R = P.to_device_order_multi(Rn, 0, "float32")
    X = torch.zeros_like(R)
    P.mg32_block(R, X, ordering=ordering)
    got = P.to_numpy_multi(X).astype(np.float32)
    for k in range(s):
        want = h.mg32(Rn[:, k])
        assert np.array_equal(got[:, k], want), k
        # and the single-vector GPU V-cycle on the same column
        z = P.zeros32()
        P.mg32(R[:, k].contiguous(), z, ordering=ordering)
        assert torch.equal(z, X[:, k].contiguous()), k
    del P


def test_cg_block_trajectories(hp):
    import torch
    dims, s = (32, 48, 64), 4
    say("cg block trajectories", dims, s)
    P = hp.Problem.generate(*dims, nlevels=4)
    h = O.Hierarchy(*dims)
    h.use_color_order()
    n = h.nrow(0)
    Bn = rhs(n, s)
    B = P.to_device_order_multi(Bn)
    want = [h.cg_mixed(b=Bn[:, k], max_iter=20, tol=0.0, recompute=10)["traj"] for k in range(s)]
    for it in (1, 5, 10, 11, 20):
        X = torch.zeros_like(B)
        r = P.cg_block_mixed(B, X, max_iter=it, tolerance=0.0, recompute=10)
        assert r["niters"] == it
        # the dots are reassociated (per-column partials of the block
        # kernels): measured relative drift up to 1.2e-9 before the first
        # FP64 residual recompute (iteration 10) and 1.1e-7 at iteration 20
        tol = 5e-9 if it < 10 else 3e-7
        for k in range(s):
            assert r["normr"][k] / r["normr0"][k] == pytest.approx(want[k][it], rel=tol), (it, k)
    # the block solution equals s single-vector solves to rounding
    X = torch.zeros_like(B)
    P.cg_block_mixed(B, X, max_iter=20, tolerance=0.0, recompute=10)
    for k in range(s):
        x = P.zeros()
        P.cg_mixed(B[:, k].contiguous(), x, max_iter=20, tolerance=0.0, recompute=10)
        d = float((x - X[:, k]).abs().max()) / float(x.abs().max())
        assert d <= 1e-7, (k, d)
    del P


def test_cg_block_reaches_tolerance(hp):
    import torch
    dims, s = (32, 48, 64), 8
    say("cg block to 1e-9", dims, s)
    P = hp.Problem.generate(*dims, nlevels=4)
    B = P.to_device_order_multi(rhs(P.nrow(0), s))
    X = torch.zeros_like(B)
    r = P.cg_block_mixed(B, X, max_iter=500, tolerance=1e-9, recompute=10)
    assert r["niters"] < 500
    assert all(a / b <= 1e-9 for a, b in zip(r["normr"], r["normr0"]))
    # the true residual of every system, recomputed in FP64
    Y = torch.zeros_like(B)
    P.spmv_block(X, Y)
    for k in range(s):
        res = float((B[:, k] - Y[:, k]).norm()) / float(B[:, k].norm())
        assert res <= 2e-9, (k, res)
    del P


def test_cg_block_graph_replays(hp):
    """The captured block iteration gives the eager launches' bits, across
    calls with different s (each recaptures) and eager work between them."""
    import torch
    dims = (32, 48, 64)
    say("cg block graph", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    n = P.nrow(0)
    want = {}
    P.tune("mixed_graph", 0)
    for s in (4, 8):
        B = P.to_device_order_multi(rhs(n, s))
        X = torch.zeros_like(B)
        r = P.cg_block_mixed(B, X, max_iter=13, tolerance=0.0, recompute=5)
        want[s] = (B, r, X.clone())
    P.tune("mixed_graph", 1)
    for s in (8, 4, 8):
        B, w, Xw = want[s]
        X = torch.zeros_like(B)
        r = P.cg_block_mixed(B, X, max_iter=13, tolerance=0.0, recompute=5)
        assert r["niters"] == w["niters"] and r["normr"] == w["normr"], s
        assert torch.equal(X, Xw), s
        x = P.zeros()
        P.cg_mixed(B[:, 0].contiguous(), x, max_iter=2, tolerance=0.0)
    del P


@pytest.mark.parametrize("s", [8])
def test_block_pairs_same_bits(hp, s):
    """block_pairs (slot 3m + 2's x rows moved W lanes from slot 3m's on
    affine tiles) changes no bit: the block V-cycle equals the oracle's per
    column, and 12 block CG iterations give the x of the run without it.
    nx = 256, so half of the finest tiles clear the x boundaries."""
    import torch
    dims = (256, 16, 24)
    say("block pairs", dims, s)
    P = hp.Problem.generate(*dims, nlevels=4)
    h = O.Hierarchy(*dims)
    h.use_color_order()
    n = h.nrow(0)
    Rn = rhs(n, s, 13).astype(np.float32)
    R = P.to_device_order_multi(Rn, 0, "float32")
    B = P.to_device_order_multi(rhs(n, s, 42), 0)
    X = torch.zeros_like(B)
    P.tune("block_pairs", 0)
    P.cg_block_mixed(B, X, max_iter=12, tolerance=0.0, recompute=5)
    X0 = X.clone()
    P.tune("block_pairs", 1)
    try:
        Z = torch.zeros_like(R)
        P.mg32_block(R, Z, ordering="multicolor")
        got = P.to_numpy_multi(Z).astype(np.float32)
        for k in range(s):
            assert np.array_equal(got[:, k], h.mg32(Rn[:, k])), k
        X.zero_()
        P.cg_block_mixed(B, X, max_iter=12, tolerance=0.0, recompute=5)
        assert torch.equal(X, X0)
    finally:
        P.tune("block_pairs", 0)
    del P


@pytest.mark.parametrize("s", [4, 8, 16])
def test_block_zero_guess_same_bits(hp, s):
    """The block V-cycle's first forward sweep from X = 0 without a fill
    (zero_guess), over every slot and over the live slots only (zero_live),
    gives the bits of the fill followed by the full sweep: the block V-cycle
    and 12 block CG iterations."""
    import torch
    dims = (32, 48, 64)
    say("block zero guess", dims, s)
    P = hp.Problem.generate(*dims, nlevels=4)
    n = P.nrow(0)
    R = P.to_device_order_multi(rhs(n, s, 17).astype(np.float32), 0, "float32")
    B = P.to_device_order_multi(rhs(n, s, 42), 0)
    out = {}
    try:
        # (zero_guess, zero_live): the filled sweep, the zero-x sweep over
        # every slot, and the zero-x sweep over the live slots only
        for zg, live in ((0, 0), (1, 0), (1, 8)):
            say("zero_guess", zg, "zero_live", live)
            P.tune("zero_guess", zg)
            P.tune("zero_live", live)
            Z = torch.full_like(R, float("nan"))  # the zero-x sweep must not read it
            if not zg:
                Z.zero_()
            P.mg32_block(R, Z, ordering="multicolor")
            X = torch.zeros_like(B)
            P.cg_block_mixed(B, X, max_iter=12, tolerance=0.0, recompute=5)
            out[(zg, live)] = (Z, X)
    finally:
        P.tune("zero_guess", 1)
        P.tune("zero_live", 11)
    for k in ((1, 0), (1, 8)):
        assert torch.equal(out[(0, 0)][0], out[k][0]), k
        assert torch.equal(out[(0, 0)][1], out[k][1]), k
    del P


@pytest.mark.parametrize("s", [4, 16])
def test_block_xr_rows(hp, s):
    """The block r.z and r update with four elements per thread and
    per-workgroup column sums (xr_rows) follow the grid-stride forms'
    trajectory to rounding (the sums are reassociated), and the captured
    iteration gives the eager launches' bits."""
    import torch
    dims = (32, 48, 64)
    say("block xr_rows", dims, s)
    P = hp.Problem.generate(*dims, nlevels=4)
    B = P.to_device_order_multi(rhs(P.nrow(0), s, 42), 0)
    out = {}
    try:
        for v, graph in ((0, 0), (1, 0), (1, 1)):
            P.tune("xr_rows", v)
            P.tune("mixed_graph", graph)
            X = torch.zeros_like(B)
            r = P.cg_block_mixed(B, X, max_iter=15, tolerance=0.0, recompute=5)
            out[(v, graph)] = (r, X)
    finally:
        P.tune("xr_rows", 1)
        P.tune("mixed_graph", 1)
    (r0, X0), (r1, X1), (rg, Xg) = out[(0, 0)], out[(1, 0)], out[(1, 1)]
    # the bounds of test_cg_block_trajectories after a residual recompute,
    # whose cancellation amplifies the reassociation (2.4e-8 measured at s = 4)
    for k in range(s):
        assert r1["normr"][k] == pytest.approx(r0["normr"][k], rel=3e-7), k
    assert float((X1 - X0).abs().max()) <= 1e-7 * float(X0.abs().max())
    assert rg["normr"] == r1["normr"] and torch.equal(Xg, X1)
    del P


@pytest.mark.parametrize("iters,tol", [(12, 0.0), (10, 0.0), (40, 1e-6)])
def test_block_defer_x(hp, iters, tol):
    """x's step taken with the next p update (defer_x) gives the bits of the
    step in the r update, eager and captured: runs that end between residual
    recomputes (12 iterations, recompute every 5), right after one (10), and
    at a tolerance."""
    import torch
    dims, s = (32, 48, 64), 8
    say("block defer_x", iters, tol)
    P = hp.Problem.generate(*dims, nlevels=4)
    B = P.to_device_order_multi(rhs(P.nrow(0), s, 42), 0)
    out = {}
    try:
        for defer in (0, 1):
            P.tune("defer_x", defer)
            for graph in (0, 1):
                P.tune("mixed_graph", graph)
                X = torch.zeros_like(B)
                r = P.cg_block_mixed(B, X, max_iter=iters, tolerance=tol, recompute=5)
                out[(defer, graph)] = (r["niters"], r["normr"], X)
    finally:
        P.tune("defer_x", 1)
        P.tune("mixed_graph", 1)
    ref = out[(0, 0)]
    for k, v in out.items():
        assert v[0] == ref[0] and v[1] == ref[1], k
        assert torch.equal(v[2], ref[2]), k
    del P


def test_block_resid_fused(hp):
    """The block residual (start and recomputes) in one pass (resid_fused)
    gives the X of the SpMM and residual passes, and the norms to rounding
    (column sums in another order)."""
    import torch
    dims, s = (32, 48, 64), 8
    say("block resid_fused", dims, s)
    P = hp.Problem.generate(*dims, nlevels=4)
    B = P.to_device_order_multi(rhs(P.nrow(0), s, 42), 0)
    out = {}
    try:
        for v in (0, 1):
            P.tune("resid_fused", v)
            X = torch.zeros_like(B)
            r = P.cg_block_mixed(B, X, max_iter=12, tolerance=0.0, recompute=5)
            out[v] = (r, X)
    finally:
        P.tune("resid_fused", 1)
    assert torch.equal(out[0][1], out[1][1])
    for k in range(s):
        assert out[1][0]["normr0"][k] == pytest.approx(out[0][0]["normr0"][k], rel=1e-13), k
        assert out[1][0]["normr"][k] == pytest.approx(out[0][0]["normr"][k], rel=1e-12), k
    del P


def test_block_nt_same_bits(hp):
    """Non-temporal multivector loads and stores in the block vector kernels
    (block_nt; multivectors of 2^24 or more values, here 128^3 x 8) change
    no bit of X or the norms."""
    import torch
    dims, s = (128, 128, 128), 8
    say("block nt", dims, s)
    P = hp.Problem.generate(*dims, nlevels=4)
    B = P.to_device_order_multi(rhs(P.nrow(0), s, 42), 0)
    out = {}
    try:
        for v in (0, 1):
            P.tune("block_nt", v)
            X = torch.zeros_like(B)
            r = P.cg_block_mixed(B, X, max_iter=12, tolerance=0.0, recompute=5)
            out[v] = (r["normr"], X)
    finally:
        P.tune("block_nt", 1)
    assert out[0][0] == out[1][0]
    assert torch.equal(out[0][1], out[1][1])
    del P
