"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Three kinds of pin:
  * golden vectors written by the compiled reference (tests/golden/, single
    threaded, so the reference sums dot products in index order like the
    oracle): compared bit for bit;
  * closed forms of the generated problem (nnz, b, A*1 = b, f2c, level and
    colour schedules);
  * independent dense algebra at small sizes (dense stencil matrix, dense
    triangular solves for SYMGS, a dense two-level V-cycle), so that a dropped
    term, a wrong sign or a transposed index fails a check that does not share
    the oracle's code.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import pyoracle as O
from tests.conftest import golden_level

GRIDS = {"16x16x16": (16, 16, 16), "8x16x24": (8, 16, 24)}


def dense_stencil(nx, ny, nz):
    """27-point matrix built from the stencil definition, not from the
    generator: 26 on the diagonal, -1 for every other point at Chebyshev
    distance 1."""
    n = nx * ny * nz
    iz, iy, ix = np.unravel_index(np.arange(n), (nz, ny, nx))
    d = np.maximum(np.maximum(abs(ix[:, None] - ix[None, :]), abs(iy[:, None] - iy[None, :])),
                   abs(iz[:, None] - iz[None, :]))
    A = np.where(d == 1, -1.0, 0.0)
    A[np.diag_indices(n)] = 26.0
    return A


def to_dense(m):
    n = m["nnz_row"].size
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(m["nnz_row"][i]):
            A[i, m["cols"][27 * i + j]] += m["vals"][27 * i + j]
    return A


def uniform(seed, n):
    """The splitmix64 uniform(-1,1) stream used by the reference harness."""
    out = np.empty(n)
    s = seed
    M = (1 << 64) - 1
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (2.0 / 9007199254740992.0) - 1.0
    return out


def test_splitmix_vectorised_matches_scalar():
    from conftest import splitmix_uniform
    for seed in (0, 42, 7):
        assert np.array_equal(splitmix_uniform(seed, 1000), uniform(seed, 1000))


# --------------------------------------------------------------------------
# generator and f2c
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", list(GRIDS))
def test_generator_matches_reference(name, request):
    g = request.getfixturevalue("g16" if name == "16x16x16" else "g81624")
    dims = GRIDS[name]
    h = O.Hierarchy(*dims)
    for l in range(4):
        m = h.level(l)
        ref = golden_level(g, l)
        for k in ("nnz_row", "diag_slot", "cols", "vals"):
            assert np.array_equal(m[k], ref[k]), (l, k)
        if l < 3:
            assert np.array_equal(m["f2c"], g["L%d_f2c" % l]), l
    assert np.array_equal(h.b, g["b"])
    assert np.array_equal(h.xexact, g["xexact"])


@pytest.mark.parametrize("dims", [(2, 2, 2), (4, 6, 8), (16, 16, 16), (32, 24, 16), (64, 64, 64)])
def test_generator_closed_forms(dims):
    nx, ny, nz = dims
    m = O.generate(*dims)
    assert m["nnz"] == (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)
    iz, iy, ix = np.unravel_index(np.arange(nx * ny * nz), (nz, ny, nx))
    per_axis = lambda i, n: np.where((i == 0) | (i == n - 1), 2, 3) if n > 1 else np.ones_like(i)
    expect = per_axis(ix, nx) * per_axis(iy, ny) * per_axis(iz, nz)
    assert np.array_equal(m["nnz_row"], expect)
    assert np.array_equal(m["b"], 27.0 - expect)
    # A*1 = b exactly (integers in double)
    assert np.array_equal(O.spmv(m, np.ones(nx * ny * nz)), m["b"])
    # columns ascending within each row, diagonal slot holds the row itself
    C = m["cols"].reshape(-1, 27)
    for j in range(26):
        used = m["nnz_row"] > j + 1
        assert np.all(C[used, j] < C[used, j + 1])
    rows = np.arange(nx * ny * nz)
    assert np.array_equal(C[rows, m["diag_slot"]], rows)


def test_norm_b_closed_form():
    # ||b|| values quoted by the survey of the reference's own output
    for dims, nb in [((16, 16, 16), 368.706), ((64, 64, 64), 1427.75)]:
        m = O.generate(*dims)
        assert abs(np.linalg.norm(m["b"]) - nb) < 1e-3 * nb


def test_generator_against_dense_stencil():
    dims = (4, 3, 5)
    m = O.generate(*dims)
    assert np.array_equal(to_dense(m), dense_stencil(*dims))


@pytest.mark.parametrize("dims", [(4, 4, 4), (8, 6, 4), (16, 16, 16)])
def test_f2c_closed_form(dims):
    nx, ny, nz = dims
    f = O.f2c(*dims)
    izc, iyc, ixc = np.unravel_index(np.arange(f.size), (nz // 2, ny // 2, nx // 2))
    assert np.array_equal(f, 2 * ixc + 2 * iyc * nx + 2 * izc * nx * ny)
    assert np.unique(f).size == f.size  # injective: prolongation has no conflicts


# --------------------------------------------------------------------------
# kernels against the reference's vectors (bit for bit)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", list(GRIDS))
def test_kernels_match_reference(name, request):
    g = request.getfixturevalue("g16" if name == "16x16x16" else "g81624")
    m = golden_level(g, 0)
    u1, u2, u3 = g["u1"], g["u2"], g["u3"]
    assert np.array_equal(u1, uniform(1, u1.size)[: u1.size]) if u1.size <= 4096 else True
    assert np.array_equal(O.spmv(m, u1), g["spmv_u1"])
    assert np.array_equal(O.spmv(m, g["xexact"]), g["spmv_xexact"])
    assert np.array_equal(O.symgs(m, u1, u2), g["symgs_u1_u2"])
    assert np.array_equal(O.symgs(m, u1, np.zeros_like(u1)), g["symgs_u1_0"])
    assert np.array_equal(O.symgs(m, g["b"], u3), g["symgs_b_u3"])
    f2c = g["L0_f2c"]
    assert np.array_equal(O.restriction(f2c, u1, O.spmv(m, u2)), g["restrict_u1_u2"])
    assert np.array_equal(O.prolongation(f2c, u3[: f2c.size], u2), g["prolong_u3_u2"])
    assert O.dot(u1, u2) == g["dot"][0]
    assert O.dot(u1, u1) == g["dot"][1]
    assert np.array_equal(O.waxpby(1.0, u1, 0.7, u2), g["waxpby_a1"])
    assert np.array_equal(O.waxpby(0.3, u1, 1.0, u2), g["waxpby_b1"])
    assert np.array_equal(O.waxpby(0.3, u1, -0.6, u2), g["waxpby_gen"])


@pytest.mark.parametrize("name", list(GRIDS))
def test_mg_matches_reference(name, request):
    g = request.getfixturevalue("g16" if name == "16x16x16" else "g81624")
    h = O.Hierarchy(*GRIDS[name])
    assert np.array_equal(h.mg(g["u1"]), g["mg_u1"])
    assert np.array_equal(h.mg(g["b"]), g["mg_b"])


# --------------------------------------------------------------------------
# kernels against independent dense algebra
# --------------------------------------------------------------------------
def test_spmv_dense():
    dims = (4, 5, 3)
    m = O.generate(*dims)
    A = dense_stencil(*dims)
    x = uniform(7, A.shape[0])
    y = O.spmv(m, x)
    bound = np.abs(A) @ np.abs(x)
    assert np.all(np.abs(y - A @ x) <= 1e-14 * bound)


def symgs_dense(A, r, x0):
    """x' solves (D+L) x' = r - U x0, then x solves (D+U) x = r - L x'
    (equal to tools/symgs.m of the reference when x0 = 0)."""
    L = np.tril(A, -1)
    U = np.triu(A, 1)
    D = np.diag(np.diag(A))
    xh = sla.solve_triangular(D + L, r - U @ x0, lower=True)
    return sla.solve_triangular(D + U, r - L @ xh, lower=False)


@pytest.mark.parametrize("dims", [(4, 4, 4), (3, 5, 2)])
def test_symgs_dense(dims):
    m = O.generate(*dims)
    A = dense_stencil(*dims)
    n = A.shape[0]
    r, x0 = uniform(11, n), uniform(12, n)
    got = O.symgs(m, r, x0)
    assert np.allclose(got, symgs_dense(A, r, x0), rtol=1e-13, atol=1e-14)
    got0 = O.symgs(m, r, np.zeros(n))
    assert np.allclose(got0, symgs_dense(A, r, np.zeros(n)), rtol=1e-13, atol=1e-14)
    # fixed point: SYMGS(b, xexact) = xexact
    assert np.allclose(O.symgs(m, m["b"], m["xexact"]), 1.0, rtol=0, atol=1e-14)


def test_mg_dense_two_level():
    dims = (4, 4, 4)
    h = O.Hierarchy(*dims, nlevels=2)
    A = dense_stencil(*dims)
    Ac = dense_stencil(2, 2, 2)
    f2c = h.level(0)["f2c"]
    n = A.shape[0]
    r = uniform(21, n)
    x1 = symgs_dense(A, r, np.zeros(n))
    rc = (r - A @ x1)[f2c]
    xc = symgs_dense(Ac, rc, np.zeros(rc.size))
    x2 = x1.copy()
    x2[f2c] += xc
    x = symgs_dense(A, r, x2)
    assert np.allclose(h.mg(r), x, rtol=1e-12, atol=1e-14)


def test_dot_waxpby_exact_sums():
    x, y = uniform(31, 5000), uniform(32, 5000)
    assert abs(O.dot(x, y) - math.fsum(x * y)) <= 1e-14 * np.sum(np.abs(x * y))
    assert np.array_equal(O.waxpby(1.0, x, 0.5, y), x + 0.5 * y)
    assert np.array_equal(O.waxpby(2.0, x, 1.0, y), 2.0 * x + y)
    assert np.array_equal(O.waxpby(2.0, x, 3.0, y), 2.0 * x + 3.0 * y)


# --------------------------------------------------------------------------
# CG, TestCG, TestSymmetry, TestNorms, flop model
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name,dims", [("16x16x16", (16, 16, 16)), ("8x16x24", (8, 16, 24)),
                                       ("32x32x32", (32, 32, 32))])
def test_cg_trajectory_matches_reference(name, dims, request):
    g = request.getfixturevalue({"16x16x16": "g16", "8x16x24": "g81624", "32x32x32": "g32"}[name])
    h = O.Hierarchy(*dims)
    res = h.cg(max_iter=50, tol=0.0, precond=True)
    assert np.array_equal(res["traj"], g["cg_traj_prec"])
    assert res["normr"] == g["cg_normr50"][0] and res["normr0"] == g["cg_normr50"][1]
    if "cg_x50" in g:
        assert np.array_equal(res["x"], g["cg_x50"])
    res = h.cg(max_iter=50, tol=0.0, precond=False)
    assert np.array_equal(res["traj"], g["cg_traj_noprec"])


def test_cg_residual_monotone_and_solution():
    h = O.Hierarchy(16, 16, 16)
    res = h.cg(max_iter=20, tol=0.0)
    t = res["traj"]
    assert np.all(np.diff(t) < 0)
    assert np.max(np.abs(res["x"] - 1.0)) < 1e-9


@pytest.mark.parametrize("name,dims", [("16x16x16", (16, 16, 16)), ("8x16x24", (8, 16, 24)),
                                       ("32x32x32", (32, 32, 32))])
def test_testcg_matches_reference(name, dims, request):
    g = request.getfixturevalue({"16x16x16": "g16", "8x16x24": "g81624", "32x32x32": "g32"}[name])
    h = O.Hierarchy(*dims)
    r = h.testcg()
    assert [r["count_pass"], r["count_fail"], r["niters_max_no_prec"],
            r["niters_max_prec"]] == list(g["testcg"])
    assert r["normr"] == g["testcg_normr"][0]
    # expectations of TestCG.cpp:87-88
    assert r["niters_max_no_prec"] <= 12 and r["niters_max_prec"] <= 2


@pytest.mark.parametrize("name", list(GRIDS))
def test_testsymmetry_matches_reference(name, request):
    g = request.getfixturevalue("g16" if name == "16x16x16" else "g81624")
    h = O.Hierarchy(*GRIDS[name])
    r = h.testsymmetry(g["sym_x"], g["sym_y"])
    assert r["depsym_spmv"] == g["testsymmetry"][0]
    assert r["depsym_mg"] == g["testsymmetry"][1]
    assert r["count_fail"] == 0 == g["testsymmetry"][2]
    assert r["residuals"] == (0.0, 0.0)


def test_testnorms_matches_reference(g16):
    r = O.testnorms([1.0e-7, 1.1e-7, 0.9e-7, 1.05e-7, 1.0e-7])
    assert (r["mean"], r["variance"], float(r["passed"])) == tuple(g16["testnorms"])
    assert not O.testnorms([0.0, 1.0])["passed"]


@pytest.mark.parametrize("name,dims", [("16x16x16", (16, 16, 16)), ("32x32x32", (32, 32, 32))])
def test_flop_model_matches_reference(name, dims, request):
    g = request.getfixturevalue({"16x16x16": "g16", "32x32x32": "g32"}[name])
    h = O.Hierarchy(*dims)
    nrow = [h.nrow(l) for l in range(4)]
    nnz = [h.nnz(l) for l in range(4)]
    # the harness reports 3 sets, refMaxIters 50, optMaxIters 52, times 1 s
    m = O.hpcg_model(nrow, nnz, 3, 50, 52)
    for key, gk in [("ddot", "model_ddot"), ("waxpby", "model_waxpby"), ("spmv", "model_spmv"),
                    ("mg", "model_mg"), ("total", "model_total"),
                    ("total_ref", "model_total_ref")]:
        assert abs(m[key] - g[gk][0]) <= 1e-5 * g[gk][0], key
    # times[0] = 1 s, so the reported B/W is bytes / 1e9
    assert abs(m["reads"] / 1e9 - g["model_read_bw"][0]) <= 1e-5 * g["model_read_bw"][0]
    assert abs(m["writes"] / 1e9 - g["model_write_bw"][0]) <= 1e-5 * g["model_write_bw"][0]


# --------------------------------------------------------------------------
# schedules
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [(16, 16, 16), (8, 16, 24), (5, 3, 4), (32, 48, 64)])
def test_levels_closed_form(dims):
    nx, ny, nz = dims
    m = O.generate(*dims)
    lev, nl = O.levels(m)
    iz, iy, ix = np.unravel_index(np.arange(nx * ny * nz), (nz, ny, nx))
    assert np.array_equal(lev, ix + 2 * iy + 4 * iz)
    assert nl == (nx - 1) + 2 * (ny - 1) + 4 * (nz - 1) + 1


@pytest.mark.parametrize("dims", [(16, 16, 16), (8, 16, 24), (32, 48, 64)])
def test_colors_closed_form(dims):
    nx, ny, nz = dims
    m = O.generate(*dims)
    col, nc = O.colors(m)
    iz, iy, ix = np.unravel_index(np.arange(nx * ny * nz), (nz, ny, nx))
    parity = (ix % 2) + 2 * (iy % 2) + 4 * (iz % 2)
    assert nc == 8
    # same partition as the parity classes (colour labels are first-fit order)
    pairs = set(zip(col.tolist(), parity.tolist()))
    assert len(pairs) == 8


def brute_force_greedy(m):
    n = m["nnz_row"].size
    C = m["cols"].reshape(-1, 27)
    color = [0] * n
    for i in range(1, n):
        used = {color[c] for c in C[i, : m["nnz_row"][i]] if c < i}
        k = 0
        while k in used:
            k += 1
        color[i] = k
    return np.array(color)


@pytest.mark.parametrize("dims", [(4, 4, 4), (3, 5, 6)])
def test_schedules_brute_force(dims):
    m = O.generate(*dims)
    C = m["cols"].reshape(-1, 27)
    lev, _ = O.levels(m)
    col, _ = O.colors(m)
    assert np.array_equal(col, brute_force_greedy(m))
    for i in range(C.shape[0]):
        nb = [c for c in C[i, : m["nnz_row"][i]] if c != i]
        assert all(col[c] != col[i] for c in nb)          # independent sets
        assert all(lev[c] != lev[i] for c in nb)          # no neighbours share a level
        assert all(lev[c] < lev[i] for c in nb if c < i)  # DAG order


def test_level_order_reproduces_natural_sweep(g16):
    m = golden_level(g16, 0)
    lev, _ = O.levels(m)
    order = np.lexsort((np.arange(lev.size), lev)).astype(np.int32)
    # any order that respects the level sets gives the natural sweep bit for bit
    assert np.array_equal(O.symgs(m, g16["u1"], g16["u2"], order), g16["symgs_u1_u2"])
    rng = np.random.default_rng(0)
    shuffled = np.lexsort((rng.permutation(lev.size), lev)).astype(np.int32)
    assert np.array_equal(O.symgs(m, g16["u1"], g16["u2"], shuffled), g16["symgs_u1_u2"])


def test_color_order_is_a_different_smoother(g16):
    m = golden_level(g16, 0)
    col, _ = O.colors(m)
    x = O.symgs(m, g16["u1"], g16["u2"], O.color_order(col))
    assert not np.array_equal(x, g16["symgs_u1_u2"])
    # still a convergent symmetric smoother: colour-ordered MG keeps TestCG's counts
    h = O.Hierarchy(16, 16, 16)
    h.use_color_order()
    r = h.testcg()
    assert r["niters_max_no_prec"] <= 12 and r["niters_max_prec"] <= 2
