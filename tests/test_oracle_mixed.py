"""Pins the oracle's mixed-precision functions (FP32 V-cycle on FP16 values).

The reference has no such path, so there are no golden vectors: the FP16
rounding is checked against numpy's IEEE binary16 conversion, and the FP32
sweeps, products and V-cycle against their FP64 counterparts (which are
pinned bit for bit to the reference) and dense algebra, at FP32 accuracy.
A wrong term, sign or index shows up as an O(1) difference there.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import pyoracle as O
from tests.test_oracle import dense_stencil, symgs_dense, uniform


def test_half_round_matches_ieee():
    rng = np.random.default_rng(3)
    vals = np.concatenate([
        rng.standard_normal(20000) * 10.0 ** rng.uniform(-9, 6, 20000),
        [0.0, -0.0, 1.0, -1.0, 26.0, 65504.0, 65519.99, 65520.0, 7e4, 2.0 ** -24,
         2.0 ** -25, 3 * 2.0 ** -26, 6.1e-5, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -12]])
    for v in vals:
        want = float(np.float16(np.float32(v)).astype(np.float32))
        got = O.half_round(v)
        assert got == want or (np.isinf(got) and np.isinf(want) and np.sign(got) == np.sign(want)), v


def test_spmv32_close_to_fp64():
    m = O.generate(8, 6, 10)
    x = uniform(5, 480)
    y32 = O.spmv32(m, x.astype(np.float32))
    y64 = O.spmv(m, x.astype(np.float32).astype(np.float64))
    bound = O.spmv(dict(m, vals=np.abs(m["vals"])), np.abs(x))
    assert np.all(np.abs(y32 - y64) <= 2e-6 * bound)


@pytest.mark.parametrize("dims", [(4, 4, 4), (3, 5, 2)])
def test_symgs32_dense(dims):
    m = O.generate(*dims)
    A = dense_stencil(*dims)
    n = A.shape[0]
    r, x0 = uniform(11, n).astype(np.float32), uniform(12, n).astype(np.float32)
    got = O.symgs32(m, r, x0)
    want = symgs_dense(A, r.astype(np.float64), x0.astype(np.float64))
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
    # fixed point: SYMGS(b, 1) = 1 to FP32 accuracy
    one = O.symgs32(m, m["b"].astype(np.float32), np.ones(n, np.float32))
    assert np.max(np.abs(one - 1.0)) < 1e-6


def test_symgs32_color_order_vs_fp64():
    m = O.generate(16, 16, 16)
    col, _ = O.colors(m)
    order = O.color_order(col)
    r, x0 = uniform(21, 4096).astype(np.float32), uniform(22, 4096).astype(np.float32)
    a = O.symgs32(m, r, x0, order)
    b = O.symgs(m, r.astype(np.float64), x0.astype(np.float64), order)
    assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(b))


@pytest.mark.parametrize("color", [False, True])
def test_mg32_vs_fp64(color):
    h = O.Hierarchy(16, 16, 16)
    if color:
        h.use_color_order()
    r = uniform(31, 4096)
    a = h.mg32(r.astype(np.float32))
    b = h.mg(r.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(b))


def test_cg_mixed_converges_like_fp64():
    h = O.Hierarchy(32, 32, 32)
    h.use_color_order()
    fp64 = h.cg(max_iter=60, tol=1e-9)
    mixed = h.cg_mixed(max_iter=200, tol=1e-9, recompute=10)
    assert mixed["normr"] / mixed["normr0"] <= 1e-9
    assert mixed["niters"] <= fp64["niters"] + 3
    assert np.max(np.abs(mixed["x"] - 1.0)) < 1e-7
    # early iterations follow the FP64 trajectory to FP32 accuracy
    t64 = h.cg(max_iter=5, tol=0.0)["traj"]
    t32 = h.cg_mixed(max_iter=5, tol=0.0)["traj"]
    assert np.allclose(t32, t64, rtol=1e-4)
