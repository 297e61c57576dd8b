"""The reported workloads beside the headline one, at their full sizes.

Mixed precision at 256^3 (FP32 V-cycle on FP16 values, FP64 outer PCG, FP64
residual every 10 iterations):
  * FP32 SpMV and the FP32 V-cycle bit-exact against the oracle's FP32
    restatement (or_spmv32, or_mg32) on the 256^3 hierarchy;
  * PCG to ||r||/||r0|| <= 1e-9 from x0 = 0 within one iteration of the
    FP64 path's count, max|x - 1| < 1e-7, max|x_mixed - x_fp64| < 1e-9.
Block PCG at 256^3, s = 8 and 16 right-hand sides (splitmix64 uniform(-1, 1)
from seed 42, one stream read row-major into [n][s]):
  * every column of the FP64 block SpMV bit-exact against the oracle's SpMV;
  * every column of the block FP32 V-cycle bit-identical to the single-vector
    V-cycle on that column;
  * to 1e-9, every column's true residual ||b - A x|| / ||b|| <= 2e-9.
The fleet of 32 problems with per-axis sizes drawn from {32, 48, 64, 96,
128} by splitmix64 seed 7:
  * the fleet's V-cycle equals, problem by problem, the V-cycle of that
    problem generated alone, bit for bit;
  * 20 fleet PCG iterations: every problem's ||r||/||r0|| within 1e-9
    relative of its own single-problem run (only the dots' order differs).
"""
import sys

import numpy as np
import pytest

from conftest import splitmix_uniform
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

D256 = (256, 256, 256)


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
def P256(hp):
    import torch
    say("generate", D256)
    p = hp.Problem.generate(*D256, nlevels=4)
    yield p
    p.close()
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def H256():
    say("oracle hierarchy", D256)
    h = O.Hierarchy(*D256)
    h.use_color_order()
    return h


def test_mixed_256_kernels_vs_oracle(P256, H256):
    import torch
    n = H256.nrow(0)
    x = splitmix_uniform(3, n).astype(np.float32)
    xd = P256.to_device_order(x.astype(np.float64)).float()
    P256.prepare_mixed()
    say("spmv32 256^3")
    y = P256.zeros32()
    P256.spmv32(xd, y)
    got = P256.to_numpy(y.double()).astype(np.float32)
    assert np.array_equal(got, O.spmv32(H256.level(0), x))
    say("mg32 256^3")
    z = P256.zeros32()
    P256.mg32(xd, z, ordering="multicolor")
    got = P256.to_numpy(z.double()).astype(np.float32)
    assert np.array_equal(got, H256.mg32(x))
    del y, z, xd
    torch.cuda.empty_cache()


def test_mixed_256_to_tolerance(P256):
    import torch
    b, _, _ = P256.vectors()
    xs, its = {}, {}
    for mode in ("fp64", "mixed"):
        say("to 1e-9", mode)
        x = torch.zeros_like(b)
        if mode == "fp64":
            r = P256.cg(b, x, max_iter=500, tolerance=1e-9, graph=True)
        else:
            r = P256.cg_mixed(b, x, max_iter=500, tolerance=1e-9, recompute=10)
        assert r["normr"] / r["normr0"] <= 1e-9, mode
        assert float((x - 1.0).abs().max()) < 1e-7, mode
        xs[mode], its[mode] = x, r["niters"]
    assert abs(its["mixed"] - its["fp64"]) <= 1, its
    assert float((xs["mixed"] - xs["fp64"]).abs().max()) < 1e-9


@pytest.mark.parametrize("s", [8, 16])
def test_block_256(P256, H256, s):
    import torch
    n = P256.nrow(0)
    Bn = splitmix_uniform(42, n * s).reshape(n, s)
    B = P256.to_device_order_multi(Bn)
    say("spmv block 256^3", s)
    Y = torch.zeros_like(B)
    P256.spmv_block(B, Y)
    got = P256.to_numpy_multi(Y)
    for k in range(s):
        assert np.array_equal(got[:, k], O.spmv(H256.level(0), Bn[:, k])), k
    del got
    say("mg32 block 256^3", s)
    R = B.float()
    Z = torch.zeros_like(R)
    P256.mg32_block(R, Z, ordering="multicolor")
    for k in range(s):
        z = P256.zeros32()
        P256.mg32(R[:, k].contiguous(), z, ordering="multicolor")
        assert torch.equal(z, Z[:, k].contiguous()), k
    del R, Z
    say("block cg to 1e-9", s)
    X = torch.zeros_like(B)
    r = P256.cg_block_mixed(B, X, max_iter=500, tolerance=1e-9, recompute=10)
    assert all(a / c <= 1e-9 for a, c in zip(r["normr"], r["normr0"])), r
    P256.spmv_block(X, Y)
    for k in range(s):
        true_res = float((B[:, k] - Y[:, k]).norm() / B[:, k].norm())
        assert true_res <= 2e-9, (k, true_res)
    del B, X, Y
    torch.cuda.empty_cache()


def test_fleet_32_seed7(hp):
    import torch
    dims = hp.fleet_dims(32, seed=7)
    assert len(dims) == 32
    say("fleet", len(dims), "problems")
    F = hp.Problem.generate_fleet(dims, nlevels=4)
    off = F.fleet_offsets(0)
    assert off[-1] == sum(a * b * c for a, b, c in dims)
    r_nat = splitmix_uniform(77, int(off[-1]))
    z = F.zeros()
    F.mg(F.to_device_order(r_nat), z, ordering="multicolor")
    zf = F.to_numpy(z)
    bf, xf, _ = F.vectors()
    say("fleet cg 20")
    rf = F.cg_fleet(bf, xf, max_iter=20, tolerance=0.0, ordering="multicolor", graph=True)
    for q, d in enumerate(dims):
        say("alone", q, d)
        Pq = hp.Problem.generate(*d, nlevels=4)
        zq = Pq.zeros()
        Pq.mg(Pq.to_device_order(r_nat[off[q]:off[q + 1]]), zq, ordering="multicolor")
        assert np.array_equal(Pq.to_numpy(zq), zf[off[q]:off[q + 1]]), (q, d)
        bq, xq, _ = Pq.vectors()
        rq = Pq.cg(bq, xq, max_iter=20, tolerance=0.0, ordering="multicolor")
        assert rf["niters"][q] == 20
        assert rf["normr0"][q] == pytest.approx(rq["normr0"], rel=1e-14)
        want = rq["normr"] / rq["normr0"]
        assert rf["normr"][q] / rf["normr0"][q] == pytest.approx(want, rel=1e-9), (q, d)
        Pq.close()
    F.close()
    torch.cuda.empty_cache()
