"""The reference's validation phase run against the device path.

Mirrors TestCG.cpp:52-123, TestSymmetry.cpp:60-138 and TestNorms.cpp:31-46:
the same steps, thresholds and pass/fail counting, with every Compute* call
going to libhpcg_amd.  Host-side vector edits (ScaleVectorValue,
FillRandomVector) are done on natural-order host copies and moved to the
device through the problem's residency helpers, as the reference's inline
helpers touch host memory.
"""
import ctypes
import sys

import numpy as np

DBL_EPSILON = sys.float_info.epsilon

_libc = None


def fill_random(n):
    """FillRandomVector (Vector.hpp:84 of the reference): rand()/RAND_MAX + 1."""
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL("libc.so.6")
    out = np.empty(n)
    for i in range(n):
        out[i] = _libc.rand() / 2147483647.0 + 1.0
    return out


def srand(seed):
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL("libc.so.6")
    _libc.srand(seed)


def test_cg(P, b, ordering="multicolor"):
    """TestCG.cpp:52-123.  b is the device-order right-hand side (restored
    on exit).  Returns the reference's TestCGData fields."""
    n = P.nrow(0)
    orig_diag = P.copy_diagonal()
    orig_b = b.clone()
    # :68-78 exaggerate the diagonal: global row < 9 -> (gid+2)*1e6, else 1e6
    scale = np.full(n, 1.0e6)
    gid = np.arange(min(9, n))
    scale[gid] = (gid + 2) * 1.0e6
    diag_nat = P.to_numpy(orig_diag)
    b_nat = P.to_numpy(b)
    P.replace_diagonal(P.to_device_order(diag_nat * scale))
    b.copy_(P.to_device_order(b_nat * scale))
    res = dict(count_pass=0, count_fail=0, expected_niters_no_prec=12, expected_niters_prec=2,
               niters_max_no_prec=0, niters_max_prec=0, normr=0.0)
    x = P.zeros()
    try:
        for k in range(2):                                   # :91-111
            expected = 2 if k == 1 else 12
            for _ in range(2):
                x.zero_()
                r = P.cg(b, x, max_iter=50, tolerance=1.0e-12, precondition=(k == 1),
                         ordering=ordering)
                if r["niters"] <= expected:
                    res["count_pass"] += 1
                else:
                    res["count_fail"] += 1
                key = "niters_max_prec" if k == 1 else "niters_max_no_prec"
                res[key] = max(res[key], r["niters"])
                res["normr"] = r["normr"]
    finally:                                                 # :114-115
        P.replace_diagonal(orig_diag)
        b.copy_(orig_b)
    return res


def test_symmetry(P, b, xexact, x_ncol=None, y_ncol=None, ordering="multicolor"):
    """TestSymmetry.cpp:60-138.  x_ncol / y_ncol are natural-order host
    vectors (default: FillRandomVector draws).  Returns depsym_spmv,
    depsym_mg, count_fail and the two A*xexact residuals against b."""
    n = P.nrow(0)
    if x_ncol is None:
        x_ncol = fill_random(n)
    if y_ncol is None:
        y_ncol = fill_random(n)
    x = P.to_device_order(x_ncol)
    y = P.to_device_order(y_ncol)
    z = P.zeros()
    ANorm = 2 * 26.0                                          # :80
    yNorm2 = P.dot(y, y)                                      # :83
    P.spmv(y, z)                                              # :84
    xtAy = P.dot(x, z)                                        # :87
    xNorm2 = P.dot(x, x)                                      # :91
    P.spmv(x, z)                                              # :92
    ytAx = P.dot(y, z)                                        # :95
    denom = (xNorm2 * ANorm * yNorm2 + yNorm2 * ANorm * xNorm2) * DBL_EPSILON
    out = dict(count_fail=0)
    out["depsym_spmv"] = abs(xtAy - ytAx) / denom             # :98
    if out["depsym_spmv"] > 1.0:
        out["count_fail"] += 1
    P.mg(y, z, ordering=ordering)                             # :105
    xtMinvy = P.dot(x, z)                                     # :108
    P.mg(x, z, ordering=ordering)                             # :112
    ytMinvx = P.dot(y, z)                                     # :115
    out["depsym_mg"] = abs(xtMinvy - ytMinvx) / denom         # :118
    if out["depsym_mg"] > 1.0:
        out["count_fail"] += 1
    res = []
    b_nat = P.to_numpy(b)
    for _ in range(2):                                        # :126-132
        P.spmv(xexact, z)
        # ComputeResidual (ComputeResidual.cpp:70-76) on host copies
        res.append(float(np.max(np.abs(b_nat - P.to_numpy(z)))))
    out["residuals"] = tuple(res)
    return out


def test_norms(values):
    """TestNorms.cpp:31-46."""
    v = np.asarray(values, np.float64)
    mean_delta = 0.0
    for i in range(v.size):
        mean_delta += v[i] - v[0]
    mean = v[0] + mean_delta / float(v.size)
    sumdiff = 0.0
    for i in range(v.size):
        sumdiff += (v[i] - mean) * (v[i] - mean)
    variance = sumdiff / float(v.size)
    return dict(mean=mean, variance=variance, passed=variance < 1.0e-6)
