"""Block sweeps (HPCGX_GS_BLOCKS) on the device against the oracle.

The device order of every level with even dimensions is the block order
(colour of the 2x2x2 block, position in the block), natural order inside a
run; a level with an odd dimension keeps the point colours.  SYMGS, the
V-cycle (FP64 and FP32) and the transfers are bit-exact against the
oracle's Gauss-Seidel in that order; the CG trajectory follows the oracle's
to 1e-9 relative (the dot products are reassociated); the block PCG equals
the single-vector path column by column.
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from test_block_order import block_order

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


_cache = {}


def problem(hp, dims):
    if dims not in _cache:
        say("generate blocks", dims)
        _cache[dims] = hp.Problem.generate(*dims, nlevels=4, gs_blocks=True)
    return _cache[dims]


def level_order(h, P, l):
    """The Gauss-Seidel order the device uses on level l (natural rows)."""
    if P.info(l)["gs_rows"] == 8:
        return block_order(*h.dims(l))
    return O.color_order(O.colors(h.level(l))[0])


_hier = {}


def hier(hp, dims):
    if dims not in _hier:
        h = O.Hierarchy(*dims)
        P = problem(hp, dims)
        for l in range(4):
            h.set_order(l, level_order(h, P, l))
        _hier[dims] = h
    return _hier[dims]


def rnd(seed, n, dtype=np.float64):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n).astype(dtype)


def device_order(P, l):
    """Natural rows listed in device order."""
    import torch
    n = P.nrow(l)
    pos = P.to_numpy(torch.arange(n, dtype=torch.float64, device="cuda"), l)
    return np.argsort(pos).astype(np.int32)


DIMS = [(32, 48, 64), (24, 24, 24), (16, 16, 16)]


@pytest.mark.parametrize("dims", DIMS)
def test_device_order_is_block_order(hp, dims):
    say("block order", dims)
    P = problem(hp, dims)
    h = O.Hierarchy(*dims)
    for l in range(4):
        info = P.info(l)
        even = all(d % 2 == 0 for d in h.dims(l))
        assert info["gs_rows"] == (8 if even else 1), l
        got = device_order(P, l)
        if even:
            assert np.array_equal(got, block_order(*h.dims(l))), l
        else:  # point colours, natural order inside each
            c = O.colors(h.level(l))[0]
            assert np.array_equal(got, O.color_order(c)), l


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("c16", [1, 0])
def test_symgs_blocks_bitexact(hp, dims, c16):
    say("symgs blocks", dims, c16)
    P = hp.Problem.generate(*dims, nlevels=4, gs_blocks=True)
    P.tune("col16", c16)
    h = hier(hp, dims)
    for l in range(4):
        m = h.level(l)
        r, x0 = rnd(10 + l, h.nrow(l)), rnd(20 + l, h.nrow(l))
        want = O.symgs(m, r, x0, level_order(h, P, l))
        x = P.to_device_order(x0, l)
        P.symgs(P.to_device_order(r, l), x, level=l, ordering="multicolor")
        assert np.array_equal(P.to_numpy(x, l), want), l
    del P


@pytest.mark.parametrize("dims", DIMS)
def test_transfers_blocks_bitexact(hp, dims):
    """The injected fine rows are position 0 of every block: eight runs,
    tile-aligned, which the fine-order transfers walk."""
    say("transfers blocks", dims)
    P = problem(hp, dims)
    h = hier(hp, dims)
    for l in range(3):
        m = h.level(l)
        f2c = m["f2c"]
        n, nc = h.nrow(l), h.nrow(l + 1)
        rf, xf, xc = rnd(30 + l, n), rnd(40 + l, n), rnd(50 + l, nc)
        want = O.restriction(f2c, rf, O.spmv(m, xf))
        rc = P.zeros(l + 1)
        P.residual_restriction(P.to_device_order(rf, l), P.to_device_order(xf, l), rc, level=l)
        assert np.array_equal(P.to_numpy(rc, l + 1), want), l
        x = P.to_device_order(xf, l)
        P.prolongation(P.to_device_order(xc, l + 1), x, level=l)
        assert np.array_equal(P.to_numpy(x, l), O.prolongation(f2c, xc, xf)), l


@pytest.mark.parametrize("dims", DIMS)
def test_mg_blocks_bitexact(hp, dims):
    say("mg blocks", dims)
    P = problem(hp, dims)
    h = hier(hp, dims)
    r = rnd(60, h.nrow(0))
    z = P.zeros()
    P.mg(P.to_device_order(r), z, ordering="multicolor")
    assert np.array_equal(P.to_numpy(z), h.mg(r))


@pytest.mark.parametrize("dims", [(32, 48, 64), (24, 24, 24)])
@pytest.mark.parametrize("c16,c8,pairs", [(1, 1, 1), (1, 1, 0), (0, 1, 1)])
def test_fp32_blocks_bitexact(hp, dims, c16, c8, pairs):
    """FP32 sweeps (FP16 value pairs, packed words, 32-bit ids) and the FP32
    V-cycle in the block order."""
    say("fp32 blocks", dims, c16, c8, pairs)
    P = hp.Problem.generate(*dims, nlevels=4, gs_blocks=True)
    P.tune("col16", c16)
    P.tune("col8", c8)
    P.tune("fp16_pairs", pairs)
    h = hier(hp, dims)
    for l in range(4):
        m = h.level(l)
        r, x0 = rnd(70 + l, h.nrow(l), np.float32), rnd(80 + l, h.nrow(l), np.float32)
        want = O.symgs32(m, r, x0, level_order(h, P, l))
        x = P.to_device_order(x0.astype(np.float64), l).float()
        P.symgs32(P.to_device_order(r.astype(np.float64), l).float(), x, level=l,
                  ordering="multicolor")
        got = P.to_numpy(x.double(), l).astype(np.float32)
        assert np.array_equal(got, want), l
    r = rnd(90, h.nrow(0), np.float32)
    z = P.zeros32()
    P.mg32(P.to_device_order(r.astype(np.float64)).float(), z, ordering="multicolor")
    assert np.array_equal(P.to_numpy(z.double()).astype(np.float32), h.mg32(r))
    del P


@pytest.mark.parametrize("s", [8, 16, 2])
def test_block_pcg_blocks_bitexact(hp, s):
    """The block V-cycle (W lanes per row for s = 8, 16; one lane per column
    and run-by-run launches for s = 2) equals the oracle's V-cycle on every
    column."""
    import torch
    dims = (32, 48, 64)
    say("mg32 block, blocks", dims, s)
    P = problem(hp, dims)
    h = hier(hp, dims)
    n = h.nrow(0)
    Rn = np.stack([rnd(100 + k, n, np.float32) for k in range(s)], axis=1)
    R = P.to_device_order_multi(Rn, 0, "float32")
    X = torch.zeros_like(R)
    P.mg32_block(R, X, ordering="multicolor")
    got = P.to_numpy_multi(X).astype(np.float32)
    for k in range(s):
        assert np.array_equal(got[:, k], h.mg32(Rn[:, k])), k


def test_cg_blocks_trajectory(hp):
    import torch
    dims = (32, 48, 64)
    say("cg blocks trajectory", dims)
    P = problem(hp, dims)
    h = hier(hp, dims)
    want = h.cg(max_iter=30, tol=0.0)["traj"]
    b, x, _ = P.vectors()
    for k in (1, 5, 15, 30):
        x.zero_()
        r = P.cg(b, x, max_iter=k, tolerance=0.0, ordering="multicolor")
        assert r["normr"] / r["normr0"] == pytest.approx(want[k], rel=1e-9), k
    # captured iterations give the eager bits
    x.zero_()
    eager = P.cg(b, x, max_iter=20, tolerance=0.0, ordering="multicolor")
    xe = x.clone()
    x.zero_()
    g = P.cg(b, x, max_iter=20, tolerance=0.0, ordering="multicolor", graph=True)
    assert g["normr"] == eager["normr"] and torch.equal(x, xe)


def test_cg_blocks_fewer_iterations(hp):
    """To the residual 50 lexicographic iterations reach, the block sweeps
    need fewer PCG iterations than the point colours (oracle: 56 and 62)."""
    dims = (32, 48, 64)
    say("cg blocks iterations", dims)
    Pp = hp.Problem.generate(*dims, nlevels=4)
    b, x, _ = Pp.vectors()
    ref = Pp.cg(b, x, max_iter=50, tolerance=0.0, ordering="lexicographic")
    tol = ref["normr"] / ref["normr0"]
    x.zero_()
    parity = Pp.cg(b, x, max_iter=200, tolerance=tol, ordering="multicolor")["niters"]
    P = problem(hp, dims)
    b, x, _ = P.vectors()
    blocks = P.cg(b, x, max_iter=200, tolerance=tol, ordering="multicolor")["niters"]
    assert blocks < parity and abs(blocks - 56) <= 1 and abs(parity - 62) <= 1, (blocks, parity)
    del Pp


def test_mixed_and_block_pcg_with_blocks(hp):
    """Mixed CG (eager and captured) and the block PCG converge on the block
    sweeps; the block PCG columns match single-vector mixed solves."""
    import torch
    dims = (32, 48, 64)
    say("mixed, block pcg with blocks", dims)
    P = problem(hp, dims)
    b, x, _ = P.vectors()
    r = P.cg_mixed(b, x, max_iter=200, tolerance=1e-9)
    assert r["normr"] / r["normr0"] <= 1e-9 and r["niters"] < 200
    s = 8
    B = torch.stack([b] * s, dim=1).contiguous()
    X = torch.zeros_like(B)
    rb = P.cg_block_mixed(B, X, max_iter=30, tolerance=0.0, recompute=10)
    x1 = P.zeros()
    P.cg_mixed(B[:, 0].contiguous(), x1, max_iter=30, tolerance=0.0, recompute=10)
    d = float((x1 - X[:, 0]).abs().max()) / float(x1.abs().max())
    assert d <= 1e-7 and rb["niters"] == 30, d


def test_host_path_blocks(hp):
    """OptimizeProblem from host matrices with grid dimensions picks the
    same block order and gives the generated problem's bits."""
    dims = (16, 16, 16)
    say("host path blocks", dims)
    h = O.Hierarchy(*dims)
    levels = []
    for l in range(4):
        m = h.level(l)
        d = dict(nnz_row=m["nnz_row"], cols=m["cols"], vals=m["vals"],
                 diag_slot=m["diag_slot"])
        d["nx"], d["ny"], d["nz"] = h.dims(l)
        if l < 3:
            d["f2c"] = m["f2c"]
        levels.append(d)
    Ph = hp.Problem.from_reference_levels(levels, gs_blocks=True)
    Pg = problem(hp, dims)
    for l in range(4):
        assert Ph.info(l)["gs_rows"] == 8
        assert np.array_equal(device_order(Ph, l), device_order(Pg, l)), l
    r = rnd(110, h.nrow(0))
    zg, zh = Pg.zeros(), Ph.zeros()
    Pg.mg(Pg.to_device_order(r), zg, ordering="multicolor")
    Ph.mg(Ph.to_device_order(r), zh, ordering="multicolor")
    assert np.array_equal(Pg.to_numpy(zg), Ph.to_numpy(zh))
    del Ph


def test_fleet_rejects_blocks(hp):
    import ctypes
    from hpcg_amd import _lib
    say("fleet rejects blocks")
    d = np.array([16, 32], np.int32)
    h = ctypes.c_void_p()
    rc = _lib.lib().hpcgx_generate_fleet(2, d.ctypes.data, d.ctypes.data, d.ctypes.data, 4,
                                         _lib.GS_BLOCKS, ctypes.byref(h))
    assert rc == 4 and not h.value  # HPCGX_EUNSUP, no handle
