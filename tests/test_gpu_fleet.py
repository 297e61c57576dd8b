"""Fleets of independent problems (hpcgx_generate_fleet / hpcgx_cg_fleet)
against the oracle, problem by problem.

A fleet's level l is the block-diagonal union of its problems' level l, so:
  * SpMV, SYMGS (both orderings, every level) and the V-cycle over the fleet
    are bit-exact against the oracle run on each problem alone;
  * fleet CG matches each problem's own CG trajectory within 1e-9 relative
    (only the dots' summation order differs), and with a tolerance every
    problem stops at its own iteration count, as the oracle's CG does;
  * iterations replayed from the captured graph give the eager bits.
Inputs: per-problem splitmix64 uniform(-1, 1) vectors, seed 5 + q.
"""
import sys

import numpy as np
import pytest

from conftest import splitmix_uniform
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

# divisible by 8 (4 levels); uneven sizes and aspect ratios
DIMS = [(16, 16, 16), (32, 16, 24), (16, 40, 16), (24, 16, 32), (8, 8, 8)]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hp():
    import torch
    assert torch.cuda.is_available()
    import hpcg_amd
    return hpcg_amd


@pytest.fixture(scope="module")
def fleet(hp):
    P = hp.Problem.generate_fleet(DIMS, nlevels=4)
    yield P
    P.close()


def hiers(color):
    hs = [O.Hierarchy(*d) for d in DIMS]
    if color:
        for h in hs:
            h.use_color_order()
    return hs


def split(v, off):
    return [v[off[q]:off[q + 1]] for q in range(len(off) - 1)]


def test_fleet_structure(fleet):
    say("fleet structure")
    hs = hiers(False)
    for l in range(4):
        off = fleet.fleet_offsets(l)
        assert list(np.diff(off)) == [h.nrow(l) for h in hs]
        info = fleet.info(l)
        assert info["nrow"] == off[-1]
        assert info["nnz"] == sum(h.nnz(l) for h in hs)
        # first-fit colours of every problem are its 8 parity classes
        assert info["ncolors"] == 8


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_fleet_spmv_symgs_bitexact(fleet, level):
    say("fleet spmv/symgs level", level)
    off = fleet.fleet_offsets(level)
    for color, ordering in ((False, "lexicographic"), (True, "multicolor")):
        hs = hiers(color)
        xn = np.concatenate([splitmix_uniform(5 + q, h.nrow(level)) for q, h in enumerate(hs)])
        rn = np.concatenate([splitmix_uniform(50 + q, h.nrow(level)) for q, h in enumerate(hs)])
        x = fleet.to_device_order(xn, level)
        y = fleet.zeros(level)
        fleet.spmv(x, y, level=level)
        got = split(fleet.to_numpy(y, level), off)
        for q, h in enumerate(hs):
            assert np.array_equal(got[q], O.spmv(h.level(level), split(xn, off)[q])), (q, level)
        r = fleet.to_device_order(rn, level)
        fleet.symgs(r, x, level=level, ordering=ordering)
        got = split(fleet.to_numpy(x, level), off)
        for q, h in enumerate(hs):
            want = split(xn, off)[q].copy()
            want = O.symgs(h.level(level), split(rn, off)[q], want,
                           order=None if not color else O.color_order(O.colors(h.level(level))[0]))
            assert np.array_equal(got[q], want), (q, level, ordering)


@pytest.mark.parametrize("ordering", ["lexicographic", "multicolor"])
def test_fleet_mg_bitexact(fleet, ordering):
    say("fleet mg", ordering)
    hs = hiers(ordering == "multicolor")
    off = fleet.fleet_offsets(0)
    rn = np.concatenate([splitmix_uniform(9 + q, h.nrow(0)) for q, h in enumerate(hs)])
    z = fleet.zeros()
    fleet.mg(fleet.to_device_order(rn), z, ordering=ordering)
    got = split(fleet.to_numpy(z), off)
    for q, h in enumerate(hs):
        assert np.array_equal(got[q], h.mg(split(rn, off)[q])), q


@pytest.mark.parametrize("ordering,precond", [("multicolor", True), ("lexicographic", True),
                                              ("multicolor", False)])
def test_fleet_cg_trajectories(fleet, ordering, precond):
    say("fleet cg", ordering, precond)
    hs = hiers(ordering == "multicolor")
    want = [h.cg(max_iter=17, tol=0.0, precond=precond)["traj"] for h in hs]
    b, x, _ = fleet.vectors()
    for its in (1, 5, 17):
        x.zero_()
        r = fleet.cg_fleet(b, x, max_iter=its, tolerance=0.0, precondition=precond,
                           ordering=ordering)
        for q, h in enumerate(hs):
            assert r["niters"][q] == its
            assert r["normr0"][q] == pytest.approx(np.linalg.norm(h.b), rel=1e-14)
            assert r["normr"][q] / r["normr0"][q] == pytest.approx(want[q][its], rel=1e-9), \
                (its, q)


def test_fleet_tolerance_per_problem(fleet):
    """Each problem stops at its own count; its x is then left alone."""
    say("fleet tolerance")
    hs = hiers(True)
    off = fleet.fleet_offsets(0)
    refs = [h.cg(max_iter=100, tol=1e-6) for h in hs]
    b, x, _ = fleet.vectors()
    r = fleet.cg_fleet(b, x, max_iter=100, tolerance=1e-6)
    assert len(set(ref["niters"] for ref in refs)) > 1  # the counts differ
    got = split(fleet.to_numpy(x), off)
    for q, ref in enumerate(refs):
        assert r["niters"][q] == ref["niters"], q
        assert r["normr"][q] / r["normr0"][q] <= 1e-6
        assert np.max(np.abs(got[q] - ref["x"])) < 1e-10, q


@pytest.mark.parametrize("tol", [0.0, 1e-6])
def test_fleet_graph_replays(fleet, tol):
    """Replayed iterations give the eager bits, call after call, with eager
    work on the same fleet between the calls."""
    import torch
    say("fleet graph", tol)
    b, x, _ = fleet.vectors()
    x.zero_()
    want = fleet.cg_fleet(b, x, max_iter=23, tolerance=tol)
    xw = x.clone()
    for rep in range(3):
        x.zero_()
        r = fleet.cg_fleet(b, x, max_iter=23, tolerance=tol, graph=True)
        assert r["niters"] == want["niters"] and r["normr"] == want["normr"], rep
        assert torch.equal(x, xw), rep
        y = fleet.zeros()
        fleet.mg(b, y, ordering="multicolor")
        x.zero_()
        fleet.cg_fleet(b, x, max_iter=2, tolerance=0.0)


@pytest.mark.parametrize("tol", [0.0, 1e-6])
def test_fleet_blocks_same_bits(fleet, tol):
    """The piece kernels with one workgroup per piece and a finishing launch
    of their own (fleet_blocks 0) give the bits of a persistent grid whose
    last workgroup finishes (fleet_blocks 3, several pieces per workgroup),
    stopping counts included."""
    import torch
    say("fleet blocks", tol)
    b, x, _ = fleet.vectors()
    out = {}
    try:
        for g in (0, 3):
            fleet.tune("fleet_blocks", g)
            x.zero_()
            r = fleet.cg_fleet(b, x, max_iter=40, tolerance=tol)
            out[g] = (r["niters"], r["normr"], x.clone())
    finally:
        fleet.tune("fleet_blocks", 0)
    assert out[0][0] == out[3][0] and out[0][1] == out[3][1]
    assert torch.equal(out[0][2], out[3][2])


def test_fleet_autotune(hp):
    say("fleet autotune")
    P = hp.Problem.generate_fleet(hp.fleet_dims(count=6, seed=7, sizes=(16, 32, 48)))
    rep = P.autotune(iterations=2)
    assert any(l.startswith("chosen fleet_blocks=") for l in rep), rep
    b, x, _ = P.vectors()
    r = P.cg_fleet(b, x, max_iter=10, graph=True)
    assert all(n == 10 for n in r["niters"])
    P.close()
