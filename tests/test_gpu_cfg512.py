"""The headline workload on the device: FP64 PCG at 512 x 512 x 640 (167.8M
rows, 4.51e9 nonzeros, 27 n > 2^32 ELL slots), 8-colour Gauss-Seidel.

The oracle does not finish at this size, so the checks are the ones the
size leaves open:
  * A 1 = b exactly (GenerateProblem_ref.cpp:178; b_i = 27 - nnz_i);
  * SpMV on a random x equals, on sampled rows (the first and last rows and
    2e5 drawn uniformly, so rows whose slots sit beyond the 2^31- and
    2^32-entry marks of the column-major ELL), the row sum written out from
    the stencil in ComputeSPMV_ref.cpp:61-70's order: neighbours in
    ascending column order, value 26 on the diagonal and -1 elsewhere,
    products added one by one;
  * xexact is a fixed point of both SYMGS orderings (to 1e-14);
  * SpMV, the colour SYMGS and the V-cycle give the same bits with 8-bit
    offsets (affine tiles on and off), 16-bit offsets and 32-bit column ids;
  * 20 PCG iterations: the residual falls at every checkpoint, the same
    bits with 8- and 16-bit offsets and from the captured graph, and within
    1e-10 relative with 32-bit ids or without the fused r.z (both sum r.z in
    another order).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (512, 512, 640)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def P():
    import torch
    assert torch.cuda.is_available()
    import hpcg_amd
    say("generate", DIMS)
    p = hpcg_amd.Problem.generate(*DIMS, nlevels=4)
    yield p
    p.close()
    torch.cuda.empty_cache()


def stencil_rows(x, rows, nx, ny, nz):
    """Row sums of the 27-point matrix at natural rows `rows`, in the
    reference's slot order (ascending columns), one product at a time."""
    ix, iy, iz = rows % nx, (rows // nx) % ny, rows // (nx * ny)
    s = np.zeros(rows.size)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                z, y, xx = iz + dz, iy + dy, ix + dx
                ok = (z >= 0) & (z < nz) & (y >= 0) & (y < ny) & (xx >= 0) & (xx < nx)
                col = np.where(ok, (z * ny + y) * nx + xx, 0)
                a = 26.0 if (dx, dy, dz) == (0, 0, 0) else -1.0
                s = np.where(ok, s + a * x[col], s)
    return s


def test_512_structure_and_ones(P):
    import torch
    info = P.info(0)
    assert info["nrow"] == 167_772_160 and info["nnz"] == 4_513_353_208
    assert info["ncolors"] == 8 and info["col16"]
    b, _, xe = P.vectors()
    y = P.zeros()
    P.spmv(xe, y)
    assert torch.equal(y, b)
    nnz = 27.0 - b
    assert float(nnz.sum()) == 4_513_353_208.0


def test_512_spmv_sampled_rows(P):
    import torch
    nx, ny, nz = DIMS
    n = nx * ny * nz
    say("spmv sampled rows")
    g = torch.Generator(device="cuda").manual_seed(512)
    xnat = torch.rand(n, device="cuda", generator=g, dtype=torch.float64) * 2 - 1
    y = P.zeros()
    P.spmv(P.to_device_order(xnat), y)
    ynat = P.to_numpy(y)
    x = xnat.cpu().numpy()
    del xnat, y
    rng = np.random.default_rng(640)
    rows = np.unique(np.concatenate([np.arange(5000), np.arange(n - 5000, n),
                                     rng.integers(0, n, 200_000)])).astype(np.int64)
    want = stencil_rows(x, rows, nx, ny, nz)
    assert np.array_equal(ynat[rows], want)


def test_512_symgs_fixed_point(P):
    b, _, xe = P.vectors()
    for ordering in ("multicolor", "lexicographic"):
        say("fixed point", ordering)
        x = xe.clone()
        P.symgs(b, x, ordering=ordering)
        assert float((x - 1.0).abs().max()) <= 1e-14, ordering


FORMATS = [("col8 affine", {}), ("col8", {"affine8": 0}), ("col16", {"col8": 0}),
           ("32-bit ids", {"col16": 0})]


def _reset(P):
    for k, v in (("col16", 1), ("col8", 1), ("affine8", 1), ("rz_fused", 1)):
        P.tune(k, v)


def test_512_formats_same_bits(P):
    import torch
    n = P.nrow(0)
    g = torch.Generator(device="cuda").manual_seed(7)
    u = torch.rand(n, device="cuda", generator=g, dtype=torch.float64) * 2 - 1
    b, _, _ = P.vectors()
    ref = None
    try:
        for name, knobs in FORMATS:
            say("format", name)
            _reset(P)
            for k, v in knobs.items():
                P.tune(k, v)
            y = P.zeros()
            P.spmv(u, y)
            x = u.clone()
            P.symgs(b, x, ordering="multicolor")
            z = P.zeros()
            P.mg(u, z, ordering="multicolor")
            if ref is None:
                ref = (y, x, z)
                continue
            assert torch.equal(y, ref[0]), name
            assert torch.equal(x, ref[1]), name
            assert torch.equal(z, ref[2]), name
    finally:
        _reset(P)


def test_512_cg(P):
    import torch
    b, x, _ = P.vectors()
    say("cg checkpoints")
    rel = []
    for k in (1, 2, 5, 10, 20):
        x.zero_()
        r = P.cg(b, x, max_iter=k, tolerance=0.0, ordering="multicolor")
        assert r["niters"] == k
        rel.append(r["normr"] / r["normr0"])
    assert all(a > c for a, c in zip(rel, rel[1:])), rel
    assert r["normr0"] == pytest.approx(float(b.norm()), rel=1e-14)
    x20 = x.clone()
    say("cg graph")
    x.zero_()
    rg = P.cg(b, x, max_iter=20, tolerance=0.0, ordering="multicolor", graph=True)
    assert rg["normr"] == r["normr"] and torch.equal(x, x20)
    try:
        say("cg col16")
        P.tune("col8", 0)
        x.zero_()
        r16 = P.cg(b, x, max_iter=20, tolerance=0.0, ordering="multicolor")
        assert r16["normr"] == r["normr"] and torch.equal(x, x20)
        _reset(P)
        for name, knob in (("32-bit ids", "col16"), ("separate r.z", "rz_fused")):
            say("cg", name)
            P.tune(knob, 0)
            x.zero_()
            ro = P.cg(b, x, max_iter=20, tolerance=0.0, ordering="multicolor")
            assert ro["normr"] == pytest.approx(r["normr"], rel=1e-10), name
            assert float((x - x20).abs().max()) <= 1e-10, name
            _reset(P)
    finally:
        _reset(P)
