"""Parity of the mixed-precision path (FP32 V-cycle on FP16 values, FP64
outer PCG) against the oracle's restatement of the same arithmetic.

Bit-exact: FP32 SpMV, both Gauss-Seidel orderings, the FP32 V-cycle (same
row operation order, no fused multiply-add, correctly rounded division).
The CG trajectory differs only through the reassociated FP64 dot products:
within 1e-9 relative over 30 iterations at 32x48x64 (as the FP64 path).
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

DIMS = [(16, 16, 16), (32, 48, 64)]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hp():
    import torch
    assert torch.cuda.is_available()
    import hpcg_amd
    return hpcg_amd


_cache = {}


def problem(hp, dims):
    if dims not in _cache:
        say("generate", dims)
        _cache[dims] = hp.Problem.generate(*dims, nlevels=4)
    return _cache[dims]


_hier = {}


def hier(dims, color):
    if (dims, color) not in _hier:
        h = O.Hierarchy(*dims)
        if color:
            h.use_color_order()
        _hier[(dims, color)] = h
    return _hier[(dims, color)]


def rnd32(seed, n):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n).astype(np.float32)


def dev32(P, v, level=0):
    return P.to_device_order(np.asarray(v, np.float64), level).float()


def host32(P, t, level=0):
    return P.to_numpy(t.double(), level).astype(np.float32)


@pytest.mark.parametrize("dims", DIMS)
def test_spmv32_bitexact(hp, dims):
    say("spmv32", dims)
    P = problem(hp, dims)
    h = hier(dims, False)
    for l in range(4):
        x = rnd32(l, h.nrow(l))
        y = P.zeros32(l)
        P.spmv32(dev32(P, x, l), y, level=l)
        assert np.array_equal(host32(P, y, l), O.spmv32(h.level(l), x)), l


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("ordering", ["lexicographic", "multicolor"])
def test_symgs32_bitexact(hp, dims, ordering):
    say("symgs32", dims, ordering)
    P = problem(hp, dims)
    h = hier(dims, False)
    for l in (0, 1, 3):
        m = h.level(l)
        r, x0 = rnd32(10 + l, h.nrow(l)), rnd32(20 + l, h.nrow(l))
        order = O.color_order(O.colors(m)[0]) if ordering == "multicolor" else None
        want = O.symgs32(m, r, x0, order)
        x = dev32(P, x0, l)
        P.symgs32(dev32(P, r, l), x, level=l, ordering=ordering)
        assert np.array_equal(host32(P, x, l), want), l


@pytest.mark.parametrize("dims", [(32, 48, 64), (64, 64, 64), (8, 8, 24)])
@pytest.mark.parametrize("c16,c8,pairs", [(1, 1, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1)])
def test_symgs32_col16(hp, dims, c16, c8, pairs):
    """Multicolour FP32 sweeps and FP32 SpMV with 8-bit column offsets and
    FP16 value pairs (the default), 8-bit offsets with the packed words, the
    packed 16-bit column / FP16 value words, and 32-bit ids: the oracle's
    bits on every level, including colour blocks whose length is not a
    multiple of 128 rows."""
    say("symgs32 col16", dims, c16, c8, pairs)
    P = hp.Problem.generate(*dims, nlevels=4)
    P.tune("col16", c16)
    P.tune("col8", c8)
    P.tune("fp16_pairs", pairs)
    h = hier(dims, False)
    for l in range(4):
        m = h.level(l)
        r, x0 = rnd32(40 + l, h.nrow(l)), rnd32(50 + l, h.nrow(l))
        want = O.symgs32(m, r, x0, O.color_order(O.colors(m)[0]))
        x = dev32(P, x0, l)
        P.symgs32(dev32(P, r, l), x, level=l, ordering="multicolor")
        assert np.array_equal(host32(P, x, l), want), (c16, c8, pairs, l)
        y = P.zeros32(l)
        P.spmv32(dev32(P, x0, l), y, level=l)
        assert np.array_equal(host32(P, y, l), O.spmv32(m, x0)), (c16, c8, pairs, l)
    del P


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("ordering", ["lexicographic", "multicolor"])
def test_mg32_bitexact(hp, dims, ordering):
    say("mg32", dims, ordering)
    P = problem(hp, dims)
    h = hier(dims, ordering == "multicolor")
    r = rnd32(30, h.nrow(0))
    z = P.zeros32()
    P.mg32(dev32(P, r), z, ordering=ordering)
    assert np.array_equal(host32(P, z), h.mg32(r))


def test_512_packed_words(hp):
    """512x512x640: 27 n = 4.5e9 ELL entries, more than one launch's 2^32
    work-items.  The packed FP16/column words must equal the 32-bit path:
    FP32 SpMV and V-cycle bit for bit, and the mixed CG's first iterations
    match the FP64-matrix ones (regression: the words were built by a
    launch over all entries)."""
    import torch
    dims = (512, 512, 640)
    say("512 packed words", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    P.prepare_mixed()
    n = P.nrow(0)
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(n, device="cuda", generator=g, dtype=torch.float32) * 2 - 1
    out = {}
    for c16 in (1, 0):
        P.tune("col16", c16)
        y, z = P.zeros32(), P.zeros32()
        P.spmv32(x, y)
        P.mg32(x, z, ordering="multicolor")
        out[c16] = (y, z)
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    b, xx, _ = P.vectors()
    res = {}
    for c16 in (1, 0):
        P.tune("col16", c16)
        xx.zero_()
        res[c16] = P.cg_mixed(b, xx, max_iter=3, tolerance=0.0, recompute=10)
    assert res[1]["normr"] == pytest.approx(res[0]["normr"], rel=1e-12)
    assert res[1]["normr"] / res[1]["normr0"] < 0.2
    # the FP64 residual recompute in one pass at this size (int64 entry
    # offsets past 2^32) against the SpMV, residual and rounding passes
    P.tune("col16", 1)
    fused = {}
    try:
        for v in (0, 1):
            P.tune("resid_fused", v)
            xx.zero_()
            r = P.cg_mixed(b, xx, max_iter=4, tolerance=0.0, recompute=2)
            fused[v] = (r["normr"], xx.clone())
    finally:
        P.tune("resid_fused", 1)
    assert torch.equal(fused[0][1], fused[1][1])
    assert fused[1][0] == pytest.approx(fused[0][0], rel=1e-12)
    del P


def test_cg_mixed_trajectory(hp):
    dims = (32, 48, 64)
    say("cg mixed trajectory", dims)
    P = problem(hp, dims)
    h = hier(dims, True)
    want = h.cg_mixed(max_iter=30, tol=0.0, recompute=10)["traj"]
    b, x, _ = P.vectors()
    for k in (1, 5, 10, 11, 20, 30):
        x.zero_()
        r = P.cg_mixed(b, x, max_iter=k, tolerance=0.0, recompute=10)
        assert r["niters"] == k
        # the FP64 residual recomputed after iteration 10, r = b - A x,
        # cancels (|r| << |b|): the reassociation-level difference in x is
        # amplified by |A||x|/|r| (about 1e-9 relative at iteration 11)
        tol = 1e-9 if k <= 10 else 1e-7
        assert r["normr"] / r["normr0"] == pytest.approx(want[k], rel=tol), k


@pytest.mark.parametrize("dims", [(32, 48, 64), (64, 64, 128)])
def test_cg_mixed_same_bits_every_format(hp, dims):
    """The mixed CG reads its matrix as FP16 value pairs with 8-bit offsets,
    as packed value + 16-bit column words, or as FP16 values with 32-bit ids:
    the same numbers in the same order, so the same bits."""
    import torch
    say("cg mixed formats", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    # r.z from the dot kernel in every format (the V-cycle's last sweep
    # leaves r.z partials only in the FP16-pair layout, summed in another order)
    P.tune("rz_fused", 0)
    b, x, _ = P.vectors()
    got = []
    for c16, pairs in ((1, 1), (1, 0), (0, 1)):
        P.tune("col16", c16)
        P.tune("fp16_pairs", pairs)
        x.zero_()
        r = P.cg_mixed(b, x, max_iter=12, tolerance=0.0, recompute=5)
        got.append((r["normr"], x.clone()))
    for k in (1, 2):
        assert got[k][0] == got[0][0], k
        assert torch.equal(got[k][1], got[0][1]), k
    del P


@pytest.mark.parametrize("tol", [0.0, 1e-6])
def test_cg_mixed_graph_replays(hp, tol):
    """The captured mixed iteration gives the eager launches' bits, call
    after call, with eager work on the same problem between the calls and a
    recapture after a tuning change (recompute every 4 iterations)."""
    import torch
    dims = (32, 48, 64)
    say("cg mixed graph", tol)
    P = hp.Problem.generate(*dims, nlevels=4)
    b, x, _ = P.vectors()
    P.tune("mixed_graph", 0)
    x.zero_()
    want = P.cg_mixed(b, x, max_iter=19, tolerance=tol, recompute=4)
    xw = x.clone()
    P.tune("mixed_graph", 1)
    for rep in range(3):
        if rep == 2:
            P.tune("stream_p", 1)  # drops the graph; same bits by construction
        x.zero_()
        r = P.cg_mixed(b, x, max_iter=19, tolerance=tol, recompute=4)
        assert r["niters"] == want["niters"] and r["normr"] == want["normr"], rep
        assert torch.equal(x, xw), rep
        z = P.zeros32()
        P.mg32(P.zeros32() + 1.0, z)
        x.zero_()
        P.cg(b, x, max_iter=2, tolerance=0.0)
    del P


@pytest.mark.parametrize("iters,tol", [(12, 0.0), (10, 0.0), (19, 1e-6)])
def test_cg_mixed_defer_x(hp, iters, tol):
    """x's step taken with the next p update (defer_x) gives the bits of the
    step in the r update, eager and captured: runs that end between
    residual recomputes (12 iterations, recompute every 5), right after one
    (10), and at a tolerance."""
    import torch
    dims = (32, 48, 64)
    say("cg mixed defer_x", iters, tol)
    P = hp.Problem.generate(*dims, nlevels=4)
    b, x, _ = P.vectors()
    out = {}
    try:
        for defer in (0, 1):
            P.tune("defer_x", defer)
            for graph in (0, 1):
                P.tune("mixed_graph", graph)
                x.zero_()
                r = P.cg_mixed(b, x, max_iter=iters, tolerance=tol, recompute=5)
                out[(defer, graph)] = (r["niters"], r["normr"], x.clone())
    finally:
        P.tune("defer_x", 1)
        P.tune("mixed_graph", 1)
    ref = out[(0, 0)]
    for k, v in out.items():
        assert v[0] == ref[0] and v[1] == ref[1], k
        assert torch.equal(v[2], ref[2]), k
    del P


def test_cg_mixed_resid_fused(hp):
    """The FP64 residual recompute in one pass (resid_fused) gives the x of
    the SpMV, residual and rounding passes, and r.r to rounding (summed in
    another order)."""
    import torch
    dims = (32, 48, 64)
    say("cg mixed resid_fused", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    b, x, _ = P.vectors()
    out = {}
    try:
        for v in (0, 1):
            P.tune("resid_fused", v)
            x.zero_()
            r = P.cg_mixed(b, x, max_iter=12, tolerance=0.0, recompute=5)
            out[v] = (r["normr"], x.clone())
    finally:
        P.tune("resid_fused", 1)
    assert torch.equal(out[0][1], out[1][1])
    assert out[1][0] == pytest.approx(out[0][0], rel=1e-12)
    del P


@pytest.mark.parametrize("dims", [(32, 48, 64), (64, 64, 64)])
def test_cg_mixed_hc_pairs(hp, dims):
    """The FP64 A p from the FP16 value pairs and 8-bit offsets (hc_pairs)
    gives the bits of the 4-byte offset/value words: the same x and r.r,
    with the affine tiles' offsets implied and with every offset loaded."""
    import torch
    say("cg mixed hc_pairs", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    assert P.info(0)["byte_tiles"] > 0
    b, x, _ = P.vectors()
    out = {}
    try:
        for aff in (1, 0):
            P.tune("affine8", aff)
            for v in (0, 1):
                P.tune("hc_pairs", v)
                x.zero_()
                r = P.cg_mixed(b, x, max_iter=12, tolerance=0.0, recompute=5)
                out[aff, v] = (r["normr"], x.clone())
    finally:
        P.tune("hc_pairs", 1)
        P.tune("affine8", 1)
    for aff in (1, 0):
        assert torch.equal(out[aff, 0][1], out[aff, 1][1]), aff
        assert out[aff, 1][0] == out[aff, 0][0], aff
    del P


def test_cg_mixed_hc_pairs_wide_tiles(hp):
    """A finest level with 8-bit and wider tiles: boundary rows of the 27-
    point problem get symmetric -1 couplings to distant rows of another
    colour (the greedy colours stay as they were, so the colour blocks stay
    on 64-row tiles, and the tiles holding the far columns leave the 8-bit
    format).  The FP64 SpMV equals the oracle's; the mixed CG with hc_pairs
    gives the bits of the packed-word path (8-bit tiles from the FP16 pairs,
    the others from the packed words) and the oracle's trajectory."""
    import torch
    dims = (16, 16, 24)
    say("hc_pairs wide tiles", dims)
    h = O.Hierarchy(*dims)
    m = h.level(0)  # live views: the oracle sees the added couplings too
    n = h.nrow(0)
    nnz, cols, vals = m["nnz_row"], m["cols"].reshape(n, 27), m["vals"].reshape(n, 27)
    color, _ = O.colors(m)
    spare = np.flatnonzero(nnz < 27)
    lo, hi = spare[spare < n // 4], spare[spare >= 3 * n // 4][::-1]
    added = 0
    for i, j in zip(lo[::7], hi[::7]):
        if color[i] == color[j] or nnz[i] >= 27 or nnz[j] >= 27:
            continue
        cols[i, nnz[i]], vals[i, nnz[i]] = j, -1.0
        cols[j, nnz[j]], vals[j, nnz[j]] = i, -1.0
        nnz[i] += 1
        nnz[j] += 1
        added += 1
    assert added >= 10
    c2, nc2 = O.colors(m)
    assert nc2 == 8 and np.array_equal(c2, color)
    levels = []
    for l in range(4):
        ml = h.level(l)
        d = dict(nnz_row=ml["nnz_row"], cols=ml["cols"], vals=ml["vals"],
                 diag_slot=ml["diag_slot"])
        if l < 3:
            d["f2c"] = ml["f2c"]
        levels.append(d)
    h.use_color_order()
    P = hp.Problem.from_reference_levels(levels)
    P.prepare_mixed()
    inf = P.info(0)
    assert inf["col16"] and 0 < inf["byte_tiles"] < n // 64, inf
    xv = np.random.default_rng(3).uniform(-1, 1, n)
    y = P.empty(0)
    P.spmv(P.to_device_order(xv), y)
    assert np.array_equal(P.to_numpy(y), O.spmv(m, xv))
    # the V-cycles (zero-x first sweeps over both tile kinds) give the
    # oracle's bits; the FP64 CG its trajectory
    rr = np.random.default_rng(4).uniform(-1, 1, n)
    z = P.zeros()
    P.mg(P.to_device_order(rr), z, ordering="multicolor")
    assert np.array_equal(P.to_numpy(z), h.mg(rr))
    z32 = P.zeros32()
    P.mg32(P.to_device_order(rr).float(), z32)
    assert np.array_equal(P.to_numpy(z32.double()).astype(np.float32),
                          h.mg32(rr.astype(np.float32)))
    b, x = P.to_device_order(h.b), P.zeros()
    want64 = h.cg(max_iter=12, tol=0.0)["traj"]
    r = P.cg(b, x, max_iter=12, tolerance=0.0, ordering="multicolor")
    assert r["normr"] / r["normr0"] == pytest.approx(want64[12], rel=1e-9)
    want = h.cg_mixed(max_iter=12, tol=0.0, recompute=5)["traj"]
    out = {}
    try:
        for v in (0, 1):
            P.tune("hc_pairs", v)
            x.zero_()
            r = P.cg_mixed(b, x, max_iter=12, tolerance=0.0, recompute=5)
            out[v] = (r["normr"], x.clone())
            assert r["normr"] / r["normr0"] == pytest.approx(want[12], rel=1e-9), v
    finally:
        P.tune("hc_pairs", 1)
    assert torch.equal(out[0][1], out[1][1])
    assert out[0][0] == out[1][0]
    del P


def test_cg_mixed_rz_fused(hp):
    """r.z from the partials of the FP32 V-cycle's last sweep (r in FP64, z
    widened) against the dot kernel over r and z: the same sum in another
    order, so the trajectories agree to rounding (1e-12 relative through
    the first residual recompute) and the fused one matches the oracle."""
    dims = (32, 48, 64)
    say("cg mixed rz fused", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    b, x, _ = P.vectors()
    h = hier(dims, True)
    want = h.cg_mixed(max_iter=10, tol=0.0, recompute=10)["traj"]
    res = {}
    for fused in (1, 0):
        P.tune("rz_fused", fused)
        x.zero_()
        res[fused] = P.cg_mixed(b, x, max_iter=10, tolerance=0.0, recompute=10)
    a, c = (res[k]["normr"] / res[k]["normr0"] for k in (1, 0))
    assert a == pytest.approx(c, rel=1e-12)
    assert a == pytest.approx(want[10], rel=1e-9)
    del P


def test_cg_mixed_reaches_tolerance(hp):
    dims = (32, 48, 64)
    say("cg mixed to 1e-9", dims)
    P = problem(hp, dims)
    h = hier(dims, True)
    ref = h.cg_mixed(max_iter=300, tol=1e-9, recompute=10)
    b, x, _ = P.vectors()
    r = P.cg_mixed(b, x, max_iter=300, tolerance=1e-9, recompute=10)
    assert r["normr"] / r["normr0"] <= 1e-9
    assert abs(r["niters"] - ref["niters"]) <= 1
    assert np.max(np.abs(P.to_numpy(x) - 1.0)) < 1e-7
    x64 = P.zeros()
    r64 = P.cg(b, x64, max_iter=300, tolerance=1e-9, ordering="multicolor")
    assert r["niters"] <= r64["niters"] + 3
    assert np.max(np.abs(P.to_numpy(x) - P.to_numpy(x64))) < 1e-7


def test_inexact_fp16_off_diagonals(hp):
    """Off-diagonal values FP16 cannot hold (-1 - k/7000, symmetric): the
    V-cycle uses their FP16 roundings as designed, but the outer A p must use
    the FP64 values, as the oracle's mixed CG does; so must the block path's
    FP64 SpMV."""
    import torch
    dims = (16, 16, 24)
    say("inexact fp16 off-diagonals", dims)
    h = O.Hierarchy(*dims)
    levels = []
    for l in range(4):
        m = h.level(l)  # live views: the oracle sees the new values too
        n = h.nrow(l)
        cols = m["cols"].reshape(n, 27)
        vals = m["vals"].reshape(n, 27)
        rows = np.repeat(np.arange(n), 27).reshape(n, 27)
        off = (cols != rows) & (np.arange(27)[None, :] < m["nnz_row"][:, None])
        vals[off] = -1.0 - ((rows[off] + cols[off]) % 7 + 1) / 7000.0
        assert not np.all(vals[off].astype(np.float16).astype(np.float64) == vals[off])
        d = dict(nnz_row=m["nnz_row"], cols=m["cols"], vals=m["vals"],
                 diag_slot=m["diag_slot"])
        if l < 3:
            d["f2c"] = m["f2c"]
        levels.append(d)
    h.use_color_order()
    P = hp.Problem.from_reference_levels(levels)
    P.prepare_mixed()
    want = h.cg_mixed(max_iter=10, tol=0.0, recompute=10)["traj"]
    b, x = P.to_device_order(h.b), P.zeros()
    for k in (1, 4, 10):
        x.zero_()
        r = P.cg_mixed(b, x, max_iter=k, tolerance=0.0, recompute=10)
        assert r["normr"] / r["normr0"] == pytest.approx(want[k], rel=1e-9), k
    n, s = P.nrow(0), 8
    Xn = np.random.default_rng(5).uniform(-1, 1, (n, s))
    Y = P.to_device_order_multi(Xn) * 0
    P.spmv_block(P.to_device_order_multi(Xn), Y)
    got = P.to_numpy_multi(Y)
    for k in range(s):
        assert np.array_equal(got[:, k], O.spmv(h.level(0), Xn[:, k])), k
    del P


def test_replace_diagonal_reaches_fp32_copy(hp):
    say("diag32")
    P = hp.Problem.generate(8, 8, 8, nlevels=1)
    P.prepare_mixed()
    d = P.copy_diagonal()
    P.replace_diagonal(d * 1.0e6)  # beyond FP16 range: kept in FP32
    x = P.zeros32() + 1.0
    y = P.zeros32()
    P.spmv32(x, y)
    h = O.Hierarchy(8, 8, 8, nlevels=1)
    want = (h.b + 26.0e6 - 26.0).astype(np.float32)
    assert np.allclose(host32(P, y), want, rtol=1e-6)
    P.replace_diagonal(d)
    P.spmv32(x, y)
    assert np.array_equal(host32(P, y), h.b.astype(np.float32))
