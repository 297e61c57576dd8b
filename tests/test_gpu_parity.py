"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Bit-exact: schedules (level and colour of every row), SpMV, both Gauss-Seidel
orderings, restriction / prolongation, the full V-cycle, WAXPBY -- the
kernels keep the reference's per-row operation order and rounding.
Dot products sum in a fixed tree: |gpu - exact| <= 1e-14 * sum|x_i y_i| and
bit-identical from call to call.  CG trajectories differ from the oracle only
through that reassociation: normr/normr0 within 1e-9 relative over the
iterations above the rounding floor.
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

SMALL = [(16, 16, 16), (8, 16, 24), (32, 48, 64)]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


@pytest.fixture(scope="module")
def hp(torch):
    import hpcg_amd
    return hpcg_amd


_cache = {}


def problem(hp, dims, keep=True):
    key = (dims, keep)
    if key not in _cache:
        say("generate", dims)
        _cache[key] = hp.Problem.generate(*dims, nlevels=4, keep_schedule=keep)
    return _cache[key]


_hier = {}


def hier(dims, color=False):
    key = (dims, color)
    if key not in _hier:
        h = O.Hierarchy(*dims)
        if color:
            h.use_color_order()
        _hier[key] = h
    return _hier[key]


def rnd(seed, n):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


# --------------------------------------------------------------------------
# setup: generator, schedules, host path
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dims", SMALL)
def test_schedules_match_oracle(hp, dims):
    P = problem(hp, dims)
    h = hier(dims)
    for l in range(4):
        say("schedule", dims, "level", l)
        m = h.level(l)
        info = P.info(l)
        assert info["nrow"] == h.nrow(l) and info["nnz"] == h.nnz(l)
        lev, nl = O.levels(m)
        col, nc = O.colors(m)
        assert info["nlevels_sched"] == nl and info["ncolors"] == nc
        if min(h.dims(l)) >= 2:
            assert nc == 8  # the parity classes
        assert np.array_equal(P.schedule(l, "level"), lev)
        assert np.array_equal(P.schedule(l, "color"), col)


def test_host_path_rows_out_of_order(hp, torch, g16):
    """Rows whose storage is not one block in row order (row i's entries in
    block n-1-i) are gathered row by row: the same problem as the generator's."""
    say("host path, rows out of order")
    levels = []
    for l in range(4):
        d = dict(nnz_row=g16["L%d_nnz_row" % l], cols=g16["L%d_cols" % l],
                 vals=g16["L%d_vals" % l], diag_slot=g16["L%d_diag_slot" % l])
        if l < 3:
            d["f2c"] = g16["L%d_f2c" % l]
        levels.append(d)
    Ph = hp.Problem.from_reference_levels(levels, keep_schedule=True, reversed_storage=True)
    Pg = problem(hp, (16, 16, 16))
    for l in range(4):
        assert np.array_equal(Ph.schedule(l, "color"), Pg.schedule(l, "color"))
    x = Pg.to_device_order(g16["u1"])
    y = Pg.zeros()
    Ph.spmv(x, y)
    assert np.array_equal(Ph.to_numpy(y), g16["spmv_u1"])
    z = Ph.zeros()
    Ph.mg(Ph.to_device_order(g16["u1"]), z, ordering="lexicographic")
    assert np.array_equal(Ph.to_numpy(z), g16["mg_u1"])


def test_schedule_proposal_checked(hp):
    """Levels with grid dimensions propose the schedule from the grid; wrong
    dimensions of the same row count must fail the row-by-row check, and the
    Kahn traversal's schedule is used instead: the oracle's either way."""
    dims = (16, 24, 32)
    h = O.Hierarchy(*dims)
    for nxyz, tag in ((dims, "right"), ((24, 16, 32), "wrong")):
        say("schedule proposal", tag)
        levels = []
        for l in range(4):
            m = h.level(l)
            d = dict(nnz_row=m["nnz_row"], cols=m["cols"], vals=m["vals"],
                     diag_slot=m["diag_slot"], nx=nxyz[0] >> l, ny=nxyz[1] >> l,
                     nz=nxyz[2] >> l)
            if l < 3:
                d["f2c"] = m["f2c"]
            levels.append(d)
        P = hp.Problem.from_reference_levels(levels, keep_schedule=True)
        for l in range(4):
            lev, nl = O.levels(h.level(l))
            col, nc = O.colors(h.level(l))
            assert P.info(l)["nlevels_sched"] == nl, (tag, l)
            assert np.array_equal(P.schedule(l, "level"), lev), (tag, l)
            assert np.array_equal(P.schedule(l, "color"), col), (tag, l)
        del P


@pytest.mark.parametrize("dims", SMALL)
def test_problem_vectors(hp, torch, dims):
    say("vectors", dims)
    P = problem(hp, dims)
    h = hier(dims)
    b, x, xe = P.vectors()
    assert np.array_equal(P.to_numpy(b), h.b)
    assert np.all(P.to_numpy(x) == 0.0) and np.all(P.to_numpy(xe) == 1.0)


def test_host_path_equals_generated(hp, torch, g16):
    """OptimizeProblem from the reference's row-pointer layout (the golden
    matrices written by the compiled reference) builds the same problem."""
    say("host path 16^3")
    levels = []
    for l in range(4):
        d = dict(nnz_row=g16["L%d_nnz_row" % l], cols=g16["L%d_cols" % l],
                 vals=g16["L%d_vals" % l], diag_slot=g16["L%d_diag_slot" % l])
        if l < 3:
            d["f2c"] = g16["L%d_f2c" % l]
        levels.append(d)
    Ph = hp.Problem.from_reference_levels(levels, keep_schedule=True)
    Pg = problem(hp, (16, 16, 16))
    for l in range(4):
        assert np.array_equal(Ph.schedule(l, "color"), Pg.schedule(l, "color"))
        assert np.array_equal(Ph.schedule(l, "level"), Pg.schedule(l, "level"))
    x = Pg.to_device_order(g16["u1"])
    y = Pg.zeros()
    Ph.spmv(x, y)
    assert np.array_equal(Ph.to_numpy(y), g16["spmv_u1"])
    z = Ph.zeros()
    Ph.mg(Ph.to_device_order(g16["u1"]), z, ordering="lexicographic")
    assert np.array_equal(Ph.to_numpy(z), g16["mg_u1"])


# --------------------------------------------------------------------------
# kernels, bit for bit
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dims", SMALL)
def test_spmv_bitexact(hp, dims):
    say("spmv", dims)
    P = problem(hp, dims)
    h = hier(dims)
    for l in range(4):
        m = h.level(l)
        x = rnd(l + 1, h.nrow(l))
        y = P.zeros(l)
        P.spmv(P.to_device_order(x, l), y, level=l)
        assert np.array_equal(P.to_numpy(y, l), O.spmv(m, x)), l
    # A*1 = b exactly
    b, _, xe = P.vectors()
    y = P.zeros()
    P.spmv(xe, y)
    assert np.array_equal(P.to_numpy(y), h.b)


def test_spmv_golden(hp, g16):
    say("spmv golden 16^3")
    P = problem(hp, (16, 16, 16))
    y = P.zeros()
    P.spmv(P.to_device_order(g16["u1"]), y)
    assert np.array_equal(P.to_numpy(y), g16["spmv_u1"])


@pytest.mark.parametrize("dims", SMALL)
@pytest.mark.parametrize("ordering", ["lexicographic", "multicolor"])
def test_symgs_bitexact(hp, dims, ordering):
    say("symgs", dims, ordering)
    P = problem(hp, dims)
    h = hier(dims)
    for l in (0, 1, 3):
        m = h.level(l)
        n = h.nrow(l)
        r, x0 = rnd(10 + l, n), rnd(20 + l, n)
        order = None
        if ordering == "multicolor":
            order = O.color_order(O.colors(m)[0])
        want = O.symgs(m, r, x0, order)
        x = P.to_device_order(x0, l)
        P.symgs(P.to_device_order(r, l), x, level=l, ordering=ordering)
        assert np.array_equal(P.to_numpy(x, l), want), l


def device_order(torch, P, l):
    """Natural rows listed in P's device order on level l."""
    n = P.nrow(l)
    pos = P.to_numpy(torch.arange(n, dtype=torch.float64, device="cuda"), l)
    return np.argsort(pos).astype(np.int32)


@pytest.mark.parametrize("sweep", [0, 1])
def test_sweep_order(hp, torch, sweep):
    """Device rows follow the colours' sweep positions (HPCGX_SWEEP_ORDER;
    natural order inside a colour), and the multicolour sweep is the
    oracle's Gauss-Seidel in that order, bit for bit."""
    dims = (32, 48, 64)
    say("sweep order", sweep)
    P = problem(hp, dims) if sweep == 0 else hp.Problem.generate(*dims, nlevels=4, sweep_order=1)
    h = hier(dims)
    for l in range(4):
        m = h.level(l)
        order = O.color_order(O.colors(m)[0], sweep)
        assert np.array_equal(device_order(torch, P, l), order), l
        r, x0 = rnd(70 + l, h.nrow(l)), rnd(80 + l, h.nrow(l))
        x = P.to_device_order(x0, l)
        P.symgs(P.to_device_order(r, l), x, level=l, ordering="multicolor")
        assert np.array_equal(P.to_numpy(x, l), O.symgs(m, r, x0, order)), l
    with pytest.raises(hp.HpcgxError):
        hp.Problem.generate(16, 16, 16, nlevels=4, sweep_order=2)


def test_sweep_order_iterations(hp, torch):
    """To the residual 50 lexicographic iterations reach, PCG with the
    default colour order needs fewer iterations than with the id order
    (oracle at 32x48x64: 53 and 62)."""
    dims = (32, 48, 64)
    say("sweep order iterations")
    P0 = problem(hp, dims)
    P1 = hp.Problem.generate(*dims, nlevels=4, sweep_order=1)
    b, x, _ = P0.vectors()
    ref = P0.cg(b, x, max_iter=50, tolerance=0.0, ordering="lexicographic")
    tol = ref["normr"] / ref["normr0"]
    its = []
    for P in (P0, P1):
        b, x, _ = P.vectors()
        its.append(P.cg(b, x, max_iter=300, tolerance=tol, ordering="multicolor")["niters"])
    assert abs(its[0] - 53) <= 1 and abs(its[1] - 62) <= 1, its
    del P1


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 48, 64)])
def test_pair_rows_bitexact(hp, dims):
    """Colour steps with two rows per thread (16-byte matrix loads) keep the
    colour-ordered result bit for bit."""
    say("pair rows", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    P.tune("col16", 0)       # the pair kernel reads 32-bit column ids
    P.tune("pair_rows", 1)
    h = hier(dims, color=True)
    for l in range(4):
        m = h.level(l)
        n = h.nrow(l)
        r, x0 = rnd(150 + l, n), rnd(160 + l, n)
        want = O.symgs(m, r, x0, O.color_order(O.colors(m)[0]))
        x = P.to_device_order(x0, l)
        P.symgs(P.to_device_order(r, l), x, level=l, ordering="multicolor")
        assert np.array_equal(P.to_numpy(x, l), want), l
    del P


@pytest.mark.parametrize("dims", SMALL + [(64, 64, 64)])
def test_col16_layout(hp, dims):
    """Tile-aligned slots with narrow column ids: built wherever the colour
    blocks start on 64-row tiles, every tile of the 27-point problem gets
    8-bit offsets, and SpMV, SYMGS and the V-cycle give the oracle's bits
    with the 8-bit, the 16-bit and the 32-bit ids of the same layout."""
    say("col16 layout", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    h = hier(dims, color=True)
    for l in range(4):
        inf = P.info(l)
        n = h.nrow(l)
        col, nc = O.colors(h.level(l))
        aligned = n >= 64 and all(c % 64 == 0 for c in np.concatenate(
            [[0], np.cumsum(np.bincount(col, minlength=nc))]))
        assert inf["col16"] == aligned, (l, inf)
        if aligned:
            assert inf["wide_tiles"] == 0, (l, inf)
            assert inf["byte_tiles"] == (n + 63) // 64, (l, inf)  # every tile: 8-bit
    r0 = rnd(301, h.nrow(0))
    want_mg = h.mg(r0)
    for c16, c8 in ((1, 1), (1, 0), (0, 1)):
        P.tune("col16", c16)
        P.tune("col8", c8)
        for l in range(4):
            m = h.level(l)
            n = h.nrow(l)
            xv, r, x0 = rnd(310 + l, n), rnd(320 + l, n), rnd(330 + l, n)
            y = P.empty(l)
            P.spmv(P.to_device_order(xv, l), y, level=l)
            assert np.array_equal(P.to_numpy(y, l), O.spmv(m, xv)), (c16, c8, l)
            want = O.symgs(m, r, x0, O.color_order(O.colors(m)[0]))
            x = P.to_device_order(x0, l)
            P.symgs(P.to_device_order(r, l), x, level=l, ordering="multicolor")
            assert np.array_equal(P.to_numpy(x, l), want), (c16, c8, l)
        z = P.zeros()
        P.mg(P.to_device_order(r0), z, ordering="multicolor")
        assert np.array_equal(P.to_numpy(z), want_mg), (c16, c8)
    del P


def test_symgs_golden(hp, g16):
    say("symgs golden 16^3")
    P = problem(hp, (16, 16, 16))
    x = P.to_device_order(g16["u2"])
    P.symgs(P.to_device_order(g16["u1"]), x, ordering="lexicographic")
    assert np.array_equal(P.to_numpy(x), g16["symgs_u1_u2"])
    x = P.zeros()
    P.symgs(P.to_device_order(g16["u1"]), x, ordering="lexicographic")
    assert np.array_equal(P.to_numpy(x), g16["symgs_u1_0"])


@pytest.mark.parametrize("dims", SMALL)
def test_transfer_bitexact(hp, dims):
    say("restriction/prolongation", dims)
    P = problem(hp, dims)
    h = hier(dims)
    for l in range(3):
        m = h.level(l)
        f2c = m["f2c"]
        n, nc = h.nrow(l), h.nrow(l + 1)
        rf, xf, xc = rnd(30 + l, n), rnd(40 + l, n), rnd(50 + l, nc)
        axf = O.spmv(m, xf)
        want = O.restriction(f2c, rf, axf)
        rc = P.zeros(l + 1)
        P.restriction(P.to_device_order(rf, l), P.to_device_order(axf, l), rc, level=l)
        assert np.array_equal(P.to_numpy(rc, l + 1), want)
        rc2 = P.zeros(l + 1)
        P.residual_restriction(P.to_device_order(rf, l), P.to_device_order(xf, l), rc2, level=l)
        assert np.array_equal(P.to_numpy(rc2, l + 1), want)
        x = P.to_device_order(xf, l)
        P.prolongation(P.to_device_order(xc, l + 1), x, level=l)
        assert np.array_equal(P.to_numpy(x, l), O.prolongation(f2c, xc, xf))


@pytest.mark.parametrize("dims", SMALL)
@pytest.mark.parametrize("ordering", ["lexicographic", "multicolor"])
def test_mg_bitexact(hp, dims, ordering):
    say("mg", dims, ordering)
    P = problem(hp, dims)
    h = hier(dims, color=(ordering == "multicolor"))
    r = rnd(60, h.nrow(0))
    z = P.zeros()
    P.mg(P.to_device_order(r), z, ordering=ordering)
    assert np.array_equal(P.to_numpy(z), h.mg(r))


def test_mg_golden(hp, g16, g81624):
    for g, dims in [(g16, (16, 16, 16)), (g81624, (8, 16, 24))]:
        say("mg golden", dims)
        P = problem(hp, dims)
        z = P.zeros()
        P.mg(P.to_device_order(g["u1"]), z, ordering="lexicographic")
        assert np.array_equal(P.to_numpy(z), g["mg_u1"])
        P.mg(P.to_device_order(g["b"]), z, ordering="lexicographic")
        assert np.array_equal(P.to_numpy(z), g["mg_b"])


@pytest.mark.parametrize("dims", SMALL + [(64, 64, 64)])
@pytest.mark.parametrize("slab_factor", [1, 3])
def test_wavefront_bitexact(hp, dims, slab_factor):
    """Multicolour sweeps run as one dependency-tracked wavefront launch per
    direction (forced onto every level) give the colour-ordered Gauss-Seidel
    result bit for bit."""
    say("wavefront", dims, slab_factor)
    P = hp.Problem.generate(*dims, nlevels=4, band_lines=1)  # natural order in colours
    P.tune("wave", 1)
    P.tune("wave_min_rows", 1)
    P.tune("wave_slab_factor", slab_factor)
    assert all(P.info(l)["wave"] for l in range(4))
    h = hier(dims, color=True)
    for l in (0, 1, 3):
        m = h.level(l)
        n = h.nrow(l)
        r, x0 = rnd(110 + l, n), rnd(120 + l, n)
        want = O.symgs(m, r, x0, O.color_order(O.colors(m)[0]))
        x = P.to_device_order(x0, l)
        P.symgs(P.to_device_order(r, l), x, level=l, ordering="multicolor")
        P.check()
        assert np.array_equal(P.to_numpy(x, l), want), l
    r = rnd(130, h.nrow(0))
    z = P.zeros()
    P.mg(P.to_device_order(r), z, ordering="multicolor")
    P.check()
    assert np.array_equal(P.to_numpy(z), h.mg(r))
    del P


@pytest.mark.parametrize("n", [100003, (1 << 24) + 5])
def test_waxpby_bitexact(hp, torch, n):
    """All three forms, and the in-place ones, at a length with a partial
    last workgroup and at one past 2^24 elements (non-temporal loads and
    stores)."""
    say("waxpby", n)
    x, y = rnd(70, n), rnd(71, n)
    X = torch.tensor(x, device="cuda")
    Y = torch.tensor(y, device="cuda")
    W = torch.empty_like(X)
    for a, b in [(1.0, 0.7), (0.3, 1.0), (0.3, -0.6)]:
        hp.Problem.waxpby(a, X, b, Y, W)
        assert np.array_equal(W.cpu().numpy(), O.waxpby(a, x, b, y)), (a, b)
    # in-place forms used by CG.cpp:115,121,122
    Wx = X.clone()
    hp.Problem.waxpby(1.0, Wx, 0.25, Y, Wx)
    assert np.array_equal(Wx.cpu().numpy(), O.waxpby(1.0, x, 0.25, y))
    Wy = Y.clone()
    hp.Problem.waxpby(1.0, X, 0.25, Wy, Wy)
    assert np.array_equal(Wy.cpu().numpy(), O.waxpby(1.0, x, 0.25, y))


@pytest.mark.parametrize("n", [1, 255, 256, 4097, 1 << 20, 3_000_017, (1 << 24) + 3])
def test_dot(hp, torch, n):
    say("dot", n)
    P = problem(hp, (16, 16, 16))
    x, y = rnd(80, n), rnd(81, n)
    X = torch.tensor(x, device="cuda")
    Y = torch.tensor(y, device="cuda")
    v = P.dot(X, Y)
    exact = float(np.sum(x.astype(np.longdouble) * y))
    assert abs(v - exact) <= 1e-14 * float(np.sum(np.abs(x * y)))
    assert P.dot(X, Y) == v  # fixed reduction order
    vv = P.dot(X, X)
    assert abs(vv - float(np.sum(x.astype(np.longdouble) ** 2))) <= 1e-14 * float(np.sum(x * x))
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    P.dot_device(X, Y, out)
    assert out.item() == v


def test_diagonal_roundtrip(hp, torch):
    say("diagonal")
    P = hp.Problem.generate(8, 8, 8, nlevels=1)
    d = P.copy_diagonal()
    assert torch.all(d == 26.0)
    P.replace_diagonal(d * 3.0)
    x = P.to_device_order(np.ones(512))
    y = P.zeros()
    P.spmv(x, y)
    h = O.Hierarchy(8, 8, 8, nlevels=1)
    assert np.array_equal(P.to_numpy(y), h.b + 52.0)  # each row gains 2*26
    P.replace_diagonal(d)
    P.spmv(x, y)
    assert np.array_equal(P.to_numpy(y), h.b)


# --------------------------------------------------------------------------
# CG and the reference's validation oracles
# --------------------------------------------------------------------------
# At 16^3 the recursive residual falls below the rounding floor after about
# 12 iterations and reassociated dot products then move it by >1e-9; the
# 32x48x64 trajectory stays within ~1e-13 through 30 iterations.
@pytest.mark.parametrize("dims,its", [((16, 16, 16), (1, 2, 5, 10, 12)),
                                      ((32, 48, 64), (1, 5, 15, 30))])
@pytest.mark.parametrize("ordering", ["lexicographic", "multicolor"])
def test_cg_trajectory(hp, torch, dims, its, ordering):
    say("cg trajectory", dims, ordering)
    P = problem(hp, dims)
    h = hier(dims, color=(ordering == "multicolor"))
    want = h.cg(max_iter=max(its), tol=0.0)["traj"]
    b, x, _ = P.vectors()
    for k in its:
        x.zero_()
        r = P.cg(b, x, max_iter=k, tolerance=0.0, ordering=ordering)
        assert r["niters"] == k
        assert r["normr0"] == pytest.approx(np.linalg.norm(h.b), rel=1e-14)
        assert r["normr"] / r["normr0"] == pytest.approx(want[k], rel=1e-9), k


def test_spmv_dot_chunks(hp, torch):
    """The fused SpMV + dot with one workgroup per row chunk: the CG
    trajectory follows the oracle's, the same bits on every call and from
    the captured graph, and the mixed CG agrees with the persistent kernel's
    to the dot's reassociation."""
    dims = (32, 48, 64)
    say("spmv dot chunks")
    P = hp.Problem.generate(*dims, nlevels=4)
    h = hier(dims, color=True)
    want = h.cg(max_iter=20, tol=0.0)["traj"]
    b, x, _ = P.vectors()
    P.tune("spmv_dot_chunks", 0)
    ref = P.cg(b, x, max_iter=20, tolerance=0.0, ordering="multicolor")
    P.tune("spmv_dot_chunks", 1)
    runs = []
    for graph in (False, False, True, True):
        x.zero_()
        r = P.cg(b, x, max_iter=20, tolerance=0.0, ordering="multicolor", graph=graph)
        assert r["normr"] / r["normr0"] == pytest.approx(want[20], rel=1e-9), graph
        runs.append((r["normr"], x.clone()))
    for nr, xv in runs[1:]:
        assert nr == runs[0][0] and torch.equal(xv, runs[0][1])
    assert runs[0][0] == pytest.approx(ref["normr"], rel=1e-9)
    x.zero_()
    m1 = P.cg_mixed(b, x, max_iter=20, tolerance=0.0)
    P.tune("spmv_dot_chunks", 0)
    x.zero_()
    m0 = P.cg_mixed(b, x, max_iter=20, tolerance=0.0)
    # mixed trajectories under reassociated dots drift further (FP32 V-cycle,
    # residual recomputed at iteration 10): 5.5e-9 measured at iteration 20
    assert m1["normr"] == pytest.approx(m0["normr"], rel=1e-7)
    del P


@pytest.mark.parametrize("dims", [(32, 48, 64), (24, 40, 56)])
def test_rz_fused(hp, torch, dims):
    """r.z from the partials of the V-cycle's last colour sweep (24x40x56:
    colour blocks of 6720 rows, so each colour's last workgroup is partial):
    the CG trajectory follows the oracle's, the same bits on every call and
    from the captured graph, and agrees with the separate dot kernel to the
    dot's reassociation.  The lexicographic ordering keeps the dot kernel."""
    say("rz fused", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    h = hier(dims, color=True)
    want = h.cg(max_iter=20, tol=0.0)["traj"]
    b, x, _ = P.vectors()
    P.tune("rz_fused", 0)
    ref = P.cg(b, x, max_iter=20, tolerance=0.0, ordering="multicolor")
    P.tune("rz_fused", 1)
    runs = []
    for graph in (False, False, True, True):
        x.zero_()
        r = P.cg(b, x, max_iter=20, tolerance=0.0, ordering="multicolor", graph=graph)
        assert r["normr"] / r["normr0"] == pytest.approx(want[20], rel=1e-9), graph
        runs.append((r["normr"], x.clone()))
    for nr, xv in runs[1:]:
        assert nr == runs[0][0] and torch.equal(xv, runs[0][1])
    assert runs[0][0] == pytest.approx(ref["normr"], rel=1e-10)
    x.zero_()
    lex1 = P.cg(b, x, max_iter=5, tolerance=0.0, ordering="lexicographic")
    P.tune("rz_fused", 0)
    x.zero_()
    lex0 = P.cg(b, x, max_iter=5, tolerance=0.0, ordering="lexicographic")
    assert lex1["normr"] == lex0["normr"]
    del P


def test_cg_graph_with_tolerance(hp, torch):
    """With a tolerance the captured iterations stop where the eager ones do,
    with the same bits."""
    dims = (32, 48, 64)
    say("cg graph tolerance")
    P = hp.Problem.generate(*dims, nlevels=4)
    b, x, _ = P.vectors()
    want = P.cg(b, x, max_iter=100, tolerance=1e-7, ordering="multicolor")
    xw = x.clone()
    for rep in range(2):
        x.zero_()
        r = P.cg(b, x, max_iter=100, tolerance=1e-7, ordering="multicolor", graph=True)
        assert r["niters"] == want["niters"] and r["normr"] == want["normr"], rep
        assert torch.equal(x, xw), rep
    del P


@pytest.mark.parametrize("ordering,precond", [("multicolor", True), ("lexicographic", True),
                                              ("multicolor", False)])
def test_cg_graph_replays(hp, torch, ordering, precond):
    """Iterations replayed from the captured hipGraph give the eager
    launches' bits, on every call, with eager work on the same problem
    between the calls and a recapture after a tuning change."""
    dims = (32, 48, 64)
    say("cg graph", ordering, precond)
    P = hp.Problem.generate(*dims, nlevels=4)
    b, x, _ = P.vectors()
    want = {}
    for c16 in (1, 0):
        P.tune("col16", c16)
        x.zero_()
        want[c16] = (P.cg(b, x, max_iter=17, tolerance=0.0, ordering=ordering,
                          precondition=precond), x.clone())
    for rep in range(3):
        for c16 in (1, 0):
            P.tune("col16", c16)  # drops the captured graph
            for again in range(2):
                x.zero_()
                r = P.cg(b, x, max_iter=17, tolerance=0.0, ordering=ordering,
                         precondition=precond, graph=True)
                w, xw = want[c16]
                assert r["niters"] == 17 and r["normr"] == w["normr"], (rep, c16, again)
                assert torch.equal(x, xw), (rep, c16, again)
                # eager work on the same problem and vectors between replays
                y = P.zeros()
                P.mg(b, y, ordering=ordering)
                x.zero_()
                P.cg(b, x, max_iter=3, tolerance=0.0, ordering=ordering)
    del P


def test_autotune(hp, torch):
    """The autotuner times every candidate, keeps one value per knob, and the
    tuned problem still solves: bit-exact V-cycle, CG within the dots'
    reassociation of the untuned run."""
    dims = (64, 64, 64)
    say("autotune", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    b, x, _ = P.vectors()
    r0 = P.cg(b, x, max_iter=10, tolerance=0.0)
    rep = P.autotune(iterations=2)
    chosen = [l for l in rep if l.startswith("chosen ")]
    assert {l.split()[1].split("=")[0] for l in chosen} == {"col8", "spmv_dot_chunks",
                                                            "spmv_dot_blocks", "spmv_order",
                                                            "xcd_remap", "gs_block",
                                                            "gs_waves", "x_pairs"}
    assert len([l for l in rep if "ms/iteration" in l]) == 2 + 2 + 4 + 2 + 2 + 4 + 2 + 2
    x.zero_()
    r1 = P.cg(b, x, max_iter=10, tolerance=0.0)
    assert r1["normr"] == pytest.approx(r0["normr"], rel=1e-10)
    h = hier(dims, color=True)
    rr = rnd(77, P.nrow(0))
    z = P.zeros()
    P.mg(P.to_device_order(rr), z, ordering="multicolor")
    assert np.array_equal(P.to_numpy(z), h.mg(rr))
    del P


@pytest.mark.parametrize("order", [1, 0])
def test_spmv_order(hp, torch, order):
    """Either chunk order gives the oracle's SpMV bits on every level and
    the oracle's CG trajectory (only the dots' summation order changes)."""
    dims = (32, 48, 64)
    say("spmv order", order)
    P = hp.Problem.generate(*dims, nlevels=4)
    P.tune("spmv_order", order)
    h = hier(dims, color=True)
    for l in range(4):
        xv = rnd(31 + l, P.nrow(l))
        y = P.zeros(l)
        P.spmv(P.to_device_order(xv, l), y, level=l)
        assert np.array_equal(P.to_numpy(y, l), O.spmv(h.level(l), xv)), l
    want = h.cg(max_iter=12, tol=0.0)["traj"]
    b, x, _ = P.vectors()
    r = P.cg(b, x, max_iter=12, tolerance=0.0, graph=True)
    assert r["normr"] / r["normr0"] == pytest.approx(want[12], rel=1e-9)
    del P


def test_cg_tolerance_stop(hp, torch):
    say("cg tolerance stop")
    P = problem(hp, (32, 48, 64))
    h = hier((32, 48, 64), color=True)
    ref = h.cg(max_iter=100, tol=1e-6)
    b, x, _ = P.vectors()
    r = P.cg(b, x, max_iter=100, tolerance=1e-6, ordering="multicolor")
    assert r["niters"] == ref["niters"]
    assert r["normr"] / r["normr0"] <= 1e-6
    assert np.max(np.abs(P.to_numpy(x) - ref["x"])) < 1e-10


@pytest.mark.parametrize("ordering", ["lexicographic", "multicolor"])
def test_testcg(hp, torch, ordering):
    from hpcg_amd import validation as V
    dims = (16, 16, 16)
    say("TestCG", dims, ordering)
    P = problem(hp, dims)
    b, _, _ = P.vectors()
    r = V.test_cg(P, b, ordering=ordering)
    want = hier(dims, color=(ordering == "multicolor")).testcg()
    assert r["count_fail"] == 0 and r["count_pass"] == 4
    assert r["niters_max_no_prec"] == want["niters_max_no_prec"]
    assert r["niters_max_prec"] == want["niters_max_prec"]
    # diagonal and b restored
    assert np.array_equal(P.to_numpy(b), hier(dims).b)


@pytest.mark.parametrize("ordering", ["lexicographic", "multicolor"])
def test_testsymmetry(hp, torch, g16, ordering):
    from hpcg_amd import validation as V
    say("TestSymmetry", ordering)
    P = problem(hp, (16, 16, 16))
    b, _, xe = P.vectors()
    r = V.test_symmetry(P, b, xe, g16["sym_x"], g16["sym_y"], ordering=ordering)
    assert r["count_fail"] == 0
    assert r["depsym_spmv"] <= 1.0 and r["depsym_mg"] <= 1.0
    assert r["residuals"] == (0.0, 0.0)


def test_testnorms_reproducible(hp, torch):
    from hpcg_amd import validation as V
    say("TestNorms")
    P = problem(hp, (32, 48, 64))
    b, x, _ = P.vectors()
    vals = []
    for _ in range(3):
        x.zero_()
        r = P.cg(b, x, max_iter=50, tolerance=0.0, ordering="multicolor")
        vals.append(r["normr"] / r["normr0"])
    assert vals[0] == vals[1] == vals[2]
    t = V.test_norms(vals)
    assert t["passed"] and t["variance"] == 0.0


# --------------------------------------------------------------------------
# larger grids: bit-exact at 128^3, size-independent properties at 256^3
# --------------------------------------------------------------------------
def test_128_bitexact(hp, torch):
    dims = (128, 128, 128)
    say("bit-exact", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    h = O.Hierarchy(*dims)
    m = h.level(0)
    x = rnd(90, h.nrow(0))
    y = P.zeros()
    P.spmv(P.to_device_order(x), y)
    assert np.array_equal(P.to_numpy(y), O.spmv(m, x))
    z = P.zeros()
    P.mg(P.to_device_order(x), z, ordering="lexicographic")
    assert np.array_equal(P.to_numpy(z), h.mg(x))
    del P


def test_256_kernels_vs_oracle(hp):
    """256^3, kernel by kernel against the oracle on the three input vectors
    x = 1, x_i = (i mod 7)/7 and splitmix64 uniform(-1, 1) from seed 0:
    SpMV, the level-scheduled (natural order) and the 8-colour SYMGS, and
    the V-cycle in both orderings -- all bit for bit."""
    from conftest import splitmix_uniform
    dims = (256, 256, 256)
    say("256^3 kernels: oracle hierarchy", dims)
    h = O.Hierarchy(*dims)
    n = h.nrow(0)
    inputs = {"ones": np.ones(n), "mod7": (np.arange(n) % 7) / 7.0,
              "splitmix0": splitmix_uniform(0, n)}
    P = hp.Problem.generate(*dims, nlevels=4)
    L0 = h.level(0)
    b = np.array(h.b)
    for name, u in inputs.items():
        say("spmv", name)
        y = P.empty()
        P.spmv(P.to_device_order(u), y)
        assert np.array_equal(P.to_numpy(y), O.spmv(L0, u)), name
    col, _ = O.colors(L0)
    corder = O.color_order(col)
    for ordering, order in (("lexicographic", None), ("multicolor", corder)):
        for name, u in inputs.items():
            say("symgs", ordering, name)
            x = P.to_device_order(u)
            P.symgs(P.to_device_order(b), x, ordering=ordering)   # r = b, x0 = u
            assert np.array_equal(P.to_numpy(x), O.symgs(L0, b, u, order)), (ordering, name)
    for ordering in ("lexicographic", "multicolor"):
        if ordering == "multicolor":
            h.use_color_order()
        for name in ("mod7", "splitmix0"):
            say("mg", ordering, name)
            u = inputs[name]
            z = P.zeros()
            P.mg(P.to_device_order(u), z, ordering=ordering)
            assert np.array_equal(P.to_numpy(z), h.mg(u)), (ordering, name)
    del P, h


def test_256_properties(hp, torch):
    dims = (256, 256, 256)
    say("properties", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    info = P.info(0)
    assert info["nnz"] == 449_455_096 and info["ncolors"] == 8
    assert info["nlevels_sched"] == 255 + 2 * 255 + 4 * 255 + 1
    b, x, xe = P.vectors()
    y = P.zeros()
    P.spmv(xe, y)
    assert torch.equal(y, b)  # A*1 = b exactly
    for ordering in ("lexicographic", "multicolor"):
        say("fixed point", ordering)
        xx = xe.clone()
        P.symgs(b, xx, ordering=ordering)
        assert float((xx - 1.0).abs().max()) <= 1e-14
    # the 8-colour smoother converges more slowly than the lexicographic one
    # (oracle, 64^3, 50 iterations: 3.9e-10 vs 1.1e-11); at 256^3 the
    # residual after 50 iterations is about 7e-3
    say("wavefront vs banded colour steps", dims)
    Pb = hp.Problem.generate(*dims, nlevels=4, band_lines=16)  # rows in bands of 16 lines
    with pytest.raises(hp.HpcgxError):
        Pb.tune("wave", 1)                    # the banded order has no wavefront schedule
    P.tune("wave", 1)
    assert P.info(0)["wave"]
    r_nat, x_nat = rnd(140, P.nrow(0)), rnd(141, P.nrow(0))
    xw = P.to_device_order(x_nat)
    P.symgs(P.to_device_order(r_nat), xw, ordering="multicolor")
    P.check()
    P.tune("wave", 0)
    xk = Pb.to_device_order(x_nat)
    Pb.symgs(Pb.to_device_order(r_nat), xk, ordering="multicolor")
    assert np.array_equal(P.to_numpy(xw), Pb.to_numpy(xk))
    del Pb
    say("cg 256^3")
    r25 = P.cg(b, x, max_iter=25, tolerance=0.0, ordering="multicolor")
    x.zero_()
    r = P.cg(b, x, max_iter=50, tolerance=0.0, ordering="multicolor")
    assert r["niters"] == 50 and r["normr"] / r["normr0"] < 2e-2
    assert r["normr"] < r25["normr"]
    del P


# (knob, value, default): each changes how the iteration is scheduled or
# shaped, never its arithmetic
SAME_BITS_KNOBS = [("zero_guess", 0, 1), ("defer_x", 0, 1), ("x_pairs", 0, 3), ("x_pairs", 15, 3),
                   ("zero_live", 0, 11), ("zero_live", 15, 11),
                   ("gs_block", 64, 256),
                   ("gs_block", 128, 256), ("gs_block", 512, 256), ("gs_waves", 0, 1)]


def test_schedule_knobs_same_bits(hp, torch):
    """The first forward sweep from the zero x without a fill (zero_guess),
    x's step taken with the next p update (defer_x), and the finest colour
    steps' workgroup size and register cap (gs_block, gs_waves) change no
    bit: the FP64 and FP32 V-cycles equal the oracle's, and 15 CG iterations
    give the defaults' x and residual.  Colours of 6720 rows end in partial
    workgroups of every size."""
    dims = (24, 40, 56)
    say("schedule knobs", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    P.prepare_mixed()
    h = hier(dims, color=True)
    rr = rnd(91, P.nrow(0))
    want_z = h.mg(rr)
    r32 = P.to_device_order(rr).float()
    z32_want = P.zeros32()
    P.mg32(r32, z32_want)
    assert np.array_equal(P.to_numpy(z32_want.double()).astype(np.float32),
                          h.mg32(rr.astype(np.float32)))
    b, x, _ = P.vectors()
    r0 = P.cg(b, x, max_iter=15, tolerance=0.0)
    x0 = x.clone()
    for knob, value, default in SAME_BITS_KNOBS:
        say("knob", knob, value)
        P.tune(knob, value)
        try:
            z = P.zeros()
            P.mg(P.to_device_order(rr), z, ordering="multicolor")
            assert np.array_equal(P.to_numpy(z), want_z), knob
            z32 = P.zeros32()
            P.mg32(r32, z32)
            assert torch.equal(z32, z32_want), knob
            x.zero_()
            r = P.cg(b, x, max_iter=15, tolerance=0.0)
            assert r["normr"] == r0["normr"] and torch.equal(x, x0), (knob, value)
        finally:
            P.tune(knob, default)
    del P


def test_x_pairs_same_bits(hp, torch):
    """x_pairs (slot 3m + 2's x values moved one lane from slot 3m's on
    affine tiles) changes no bit: the FP64 and FP32 V-cycles equal the
    oracle's, and 12 CG and 12 mixed CG iterations give the x and residual
    of the run without it.  nx = 256, so half of the finest tiles clear the
    x boundaries and take that path."""
    dims = (256, 16, 24)
    say("x_pairs", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    P.prepare_mixed()
    h = hier(dims, color=True)
    rr = rnd(93, P.nrow(0))
    want_z = h.mg(rr)
    want_z32 = h.mg32(rr.astype(np.float32))
    r32 = P.to_device_order(rr).float()
    b, x, _ = P.vectors()
    P.tune("x_pairs", 0)
    r0 = P.cg(b, x, max_iter=12, tolerance=0.0)
    x0 = x.clone()
    x.zero_()
    m0 = P.cg_mixed(b, x, max_iter=12, tolerance=0.0, recompute=5)
    xm0 = x.clone()
    P.tune("x_pairs", 15)
    try:
        z = P.zeros()
        P.mg(P.to_device_order(rr), z, ordering="multicolor")
        assert np.array_equal(P.to_numpy(z), want_z)
        z32 = P.zeros32()
        P.mg32(r32, z32)
        assert np.array_equal(P.to_numpy(z32.double()).astype(np.float32), want_z32)
        for graph in (False, True):
            x.zero_()
            r = P.cg(b, x, max_iter=12, tolerance=0.0, graph=graph)
            assert r["normr"] == r0["normr"] and torch.equal(x, x0), graph
        x.zero_()
        m = P.cg_mixed(b, x, max_iter=12, tolerance=0.0, recompute=5)
        assert m["normr"] == m0["normr"] and torch.equal(x, xm0)
    finally:
        P.tune("x_pairs", 3)
    del P


def test_zero_live_same_bits(hp, torch):
    """zero_live (the zero-x first sweeps' loads issued for the live slots
    only, eight per chunk) changes no bit with 8-bit affine, 8-bit and
    16-bit column offsets: the FP64 and FP32 V-cycles equal the oracle's,
    and 12 CG and 12 mixed CG iterations give the x and residual of the run
    without it."""
    dims = (256, 16, 24)
    say("zero_live", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    P.prepare_mixed()
    h = hier(dims, color=True)
    rr = rnd(97, P.nrow(0))
    want_z = h.mg(rr)
    want_z32 = h.mg32(rr.astype(np.float32))
    r32 = P.to_device_order(rr).float()
    b, x, _ = P.vectors()
    try:
        for c8, a8 in ((1, 1), (1, 0), (0, 1)):
            P.tune("col8", c8)
            P.tune("affine8", a8)
            # the mixed outer iteration's bits depend on the column format,
            # so each format's runs are compared with zero_live off in it
            ref = None
            for live in (0, 15):
                say("zero_live", live, "col8", c8, "affine8", a8)
                P.tune("zero_live", live)
                z = P.zeros()
                P.mg(P.to_device_order(rr), z, ordering="multicolor")
                assert np.array_equal(P.to_numpy(z), want_z), (c8, a8, live)
                z32 = P.zeros32()
                P.mg32(r32, z32)
                assert np.array_equal(P.to_numpy(z32.double()).astype(np.float32),
                                      want_z32), (c8, a8, live)
                x.zero_()
                r = P.cg(b, x, max_iter=12, tolerance=0.0)
                xr = x.clone()
                x.zero_()
                m = P.cg_mixed(b, x, max_iter=12, tolerance=0.0, recompute=5)
                got = (r["normr"], xr, m["normr"], x.clone())
                if ref is None:
                    ref = got
                    continue
                assert got[0] == ref[0] and torch.equal(got[1], ref[1]), (c8, a8, live)
                assert got[2] == ref[2] and torch.equal(got[3], ref[3]), (c8, a8, live)
    finally:
        for k, v in (("col8", 1), ("affine8", 1), ("zero_live", 11)):
            P.tune(k, v)
    del P


def test_xr_rows(hp, torch):
    """The one-row r update (xr_rows) sums r.r in another order and changes
    nothing else: x keeps its bits in the FP64 and the mixed CG (alpha and
    beta do not use r.r), and the residual norm agrees to rounding."""
    dims = (24, 40, 56)
    say("xr_rows", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    b, x, _ = P.vectors()
    res = {}
    for v in (0, 1):
        P.tune("xr_rows", v)
        x.zero_()
        r = P.cg(b, x, max_iter=15, tolerance=0.0)
        xf = x.clone()
        x.zero_()
        m = P.cg_mixed(b, x, max_iter=15, tolerance=0.0, recompute=4)
        res[v] = (r, xf, m, x.clone())
    P.tune("xr_rows", 1)
    (r0, x0, m0, xm0), (r1, x1, m1, xm1) = res[0], res[1]
    assert torch.equal(x0, x1) and torch.equal(xm0, xm1)
    assert r1["normr"] == pytest.approx(r0["normr"], rel=1e-13)
    assert m1["normr"] == pytest.approx(m0["normr"], rel=1e-13)
    del P
