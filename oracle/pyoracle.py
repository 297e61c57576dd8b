"""ctypes wrapper over oracle/liboracle.so -- the CPU restatement of the
HPCG 3.1 reference PCG path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the hpcg_amd package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
SLOTS = 27

_lib = None


def build():
    """Compile liboracle.so (gcc, seconds)."""
    subprocess.run(["make", "-C", HERE, "-s", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        I = ctypes.c_int
        D = ctypes.c_double
        sig = {
            "or_generate": (I64, [I, I, I, P, P, P, P, P, P, P]),
            "or_f2c": (None, [I, I, I, P]),
            "or_spmv": (None, [I64, P, P, P, P, P]),
            "or_symgs": (None, [I64, P, P, P, P, P, P, P]),
            "or_restriction": (None, [I64, P, P, P, P]),
            "or_prolongation": (None, [I64, P, P, P]),
            "or_dot": (D, [I64, P, P]),
            "or_waxpby": (None, [I64, D, P, D, P, P]),
            "or_residual": (D, [I64, P, P]),
            "or_levels": (I, [I64, P, P, P]),
            "or_colors": (I, [I64, P, P, P]),
            "or_hier_create": (P, [I, I, I, I]),
            "or_hier_destroy": (None, [P]),
            "or_hier_nlevels": (I, [P]),
            "or_hier_nrow": (I64, [P, I]),
            "or_hier_nnz": (I64, [P, I]),
            "or_hier_dims": (None, [P, I, P, P, P]),
            "or_hier_nnz_row": (P, [P, I]),
            "or_hier_cols": (P, [P, I]),
            "or_hier_vals": (P, [P, I]),
            "or_hier_diag_slot": (P, [P, I]),
            "or_hier_f2c": (P, [P, I]),
            "or_hier_b": (P, [P]),
            "or_hier_x": (P, [P]),
            "or_hier_xexact": (P, [P]),
            "or_hier_set_order": (I, [P, I, P]),
            "or_mg": (I, [P, I, P, P]),
            "or_cg": (I, [P, P, P, I, D, I, P, P, P, P]),
            "or_testcg": (I, [P, P, P, P]),
            "or_testsymmetry": (I, [P, P, P, P, P, P]),
            "or_testnorms": (None, [I, P, P]),
            "or_hpcg_model": (None, [I, P, P, I, I, I, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _view(ptr, n, dtype):
    if not ptr:
        return None
    ct = {np.float64: ctypes.c_double, np.int32: ctypes.c_int32, np.uint8: ctypes.c_uint8}[dtype]
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(n,))


def generate(nx, ny, nz):
    """GenerateProblem_ref restated: dict of per-row arrays and vectors."""
    n = nx * ny * nz
    d = dict(nnz_row=np.zeros(n, np.uint8), cols=np.zeros(n * SLOTS, np.int32),
             vals=np.zeros(n * SLOTS, np.float64), diag_slot=np.zeros(n, np.uint8),
             b=np.zeros(n), x=np.zeros(n), xexact=np.zeros(n))
    d["nnz"] = lib().or_generate(nx, ny, nz, _p(d["nnz_row"]), _p(d["cols"]), _p(d["vals"]),
                                 _p(d["diag_slot"]), _p(d["b"]), _p(d["x"]), _p(d["xexact"]))
    return d


def f2c(nxf, nyf, nzf):
    out = np.zeros((nxf // 2) * (nyf // 2) * (nzf // 2), np.int32)
    lib().or_f2c(nxf, nyf, nzf, _p(out))
    return out


def spmv(m, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(m["nnz_row"].size)
    lib().or_spmv(y.size, _p(m["nnz_row"]), _p(m["cols"]), _p(m["vals"]), _p(x), _p(y))
    return y


def symgs(m, r, x, order=None):
    r = np.ascontiguousarray(r, np.float64)
    x = np.array(x, np.float64, copy=True)
    o = None if order is None else np.ascontiguousarray(order, np.int32)
    lib().or_symgs(x.size, _p(m["nnz_row"]), _p(m["cols"]), _p(m["vals"]), _p(m["diag_slot"]),
                   _p(o), _p(r), _p(x))
    return x


def restriction(f2c_, rf, axf):
    rc = np.zeros(f2c_.size)
    lib().or_restriction(f2c_.size, _p(f2c_), _p(np.ascontiguousarray(rf)),
                         _p(np.ascontiguousarray(axf)), _p(rc))
    return rc


def prolongation(f2c_, xc, xf):
    xf = np.array(xf, np.float64, copy=True)
    lib().or_prolongation(f2c_.size, _p(f2c_), _p(np.ascontiguousarray(xc)), _p(xf))
    return xf


def dot(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = x if y is x else np.ascontiguousarray(y, np.float64)
    return lib().or_dot(x.size, _p(x), _p(y))


def waxpby(alpha, x, beta, y):
    w = np.zeros(len(x))
    lib().or_waxpby(w.size, alpha, _p(np.ascontiguousarray(x, np.float64)), beta,
                    _p(np.ascontiguousarray(y, np.float64)), _p(w))
    return w


def levels(m):
    out = np.zeros(m["nnz_row"].size, np.int32)
    nl = lib().or_levels(out.size, _p(m["nnz_row"]), _p(m["cols"]), _p(out))
    return out, nl


def colors(m):
    out = np.zeros(m["nnz_row"].size, np.int32)
    nc = lib().or_colors(out.size, _p(m["nnz_row"]), _p(m["cols"]), _p(out))
    return out, nc


# Forward visiting orders of the greedy colours 0..7 (colours from 8 on
# follow in id order).  SWEEP_ORDERS[0] visits colour 0 -- on the 27-point
# grid the even points, which the restriction injects (ComputeRestriction_ref
# .cpp:51, GenerateCoarseProblem.cpp:77-79) -- last in the forward sweep, so
# the backward sweep updates it first; SWEEP_ORDERS[1] is the id order.
SWEEP_ORDERS = ((3, 5, 6, 7, 1, 2, 4, 0), (0, 1, 2, 3, 4, 5, 6, 7))


def color_order(color, sweep=0):
    """Rows in the Gauss-Seidel order of the multicolour sweep: by the
    colour's position in SWEEP_ORDERS[sweep], then by row."""
    color = np.asarray(color)
    rank = np.arange(max(int(color.max()) + 1, 8), dtype=np.int64)
    rank[list(SWEEP_ORDERS[sweep])] = np.arange(8)
    return np.lexsort((np.arange(color.size), rank[color])).astype(np.int32)


def hpcg_model(nrow, nnz, nsets, ref_iters, opt_iters):
    out = np.zeros(10)
    nrow = np.ascontiguousarray(nrow, np.int64)
    nnz = np.ascontiguousarray(nnz, np.int64)
    lib().or_hpcg_model(nrow.size, _p(nrow), _p(nnz), nsets, ref_iters, opt_iters, _p(out))
    keys = ["ddot", "waxpby", "spmv", "mg", "total", "total_ref", "reads", "writes",
            "reads_ref", "writes_ref"]
    return dict(zip(keys, out))


def testnorms(values):
    v = np.ascontiguousarray(values, np.float64)
    out = np.zeros(3)
    lib().or_testnorms(v.size, _p(v), _p(out))
    return dict(mean=out[0], variance=out[1], passed=bool(out[2]))


class Hierarchy:
    """The 4-level MG hierarchy of main.cpp:133-144, owned by the C side."""

    def __init__(self, nx, ny, nz, nlevels=4):
        self._h = lib().or_hier_create(nx, ny, nz, nlevels)
        if not self._h:
            raise ValueError("bad hierarchy dimensions %r" % ((nx, ny, nz, nlevels),))
        self.nlevels = nlevels

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_hier_destroy(self._h)
            self._h = None

    def nrow(self, l):
        return lib().or_hier_nrow(self._h, l)

    def nnz(self, l):
        return lib().or_hier_nnz(self._h, l)

    def dims(self, l):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().or_hier_dims(self._h, l, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def level(self, l):
        """Live numpy views of level l's storage (writes go to the C side)."""
        n = self.nrow(l)
        L = lib()
        d = dict(nnz_row=_view(L.or_hier_nnz_row(self._h, l), n, np.uint8),
                 cols=_view(L.or_hier_cols(self._h, l), n * SLOTS, np.int32),
                 vals=_view(L.or_hier_vals(self._h, l), n * SLOTS, np.float64),
                 diag_slot=_view(L.or_hier_diag_slot(self._h, l), n, np.uint8))
        if l + 1 < self.nlevels:
            d["f2c"] = _view(L.or_hier_f2c(self._h, l), self.nrow(l + 1), np.int32)
        return d

    @property
    def b(self):
        return _view(lib().or_hier_b(self._h), self.nrow(0), np.float64)

    @property
    def xexact(self):
        return _view(lib().or_hier_xexact(self._h), self.nrow(0), np.float64)

    def set_order(self, l, order):
        o = None if order is None else np.ascontiguousarray(order, np.int32)
        lib().or_hier_set_order(self._h, l, _p(o))

    def use_color_order(self, sweep=0):
        for l in range(self.nlevels):
            c, _ = colors(self.level(l))
            self.set_order(l, color_order(c, sweep))

    def mg(self, r, level=0):
        r = np.ascontiguousarray(r, np.float64)
        x = np.zeros(self.nrow(level))
        lib().or_mg(self._h, level, _p(r), _p(x))
        return x

    def cg(self, b=None, x=None, max_iter=50, tol=0.0, precond=True):
        b = self.b.copy() if b is None else np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.nrow(0)) if x is None else np.array(x, np.float64, copy=True)
        traj = np.zeros(max_iter + 1)
        ni, nr, nr0 = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        lib().or_cg(self._h, _p(b), _p(x), max_iter, tol, int(precond), ctypes.byref(ni),
                    ctypes.byref(nr), ctypes.byref(nr0), _p(traj))
        return dict(x=x, niters=ni.value, normr=nr.value, normr0=nr0.value,
                    traj=traj[: ni.value + 1])

    def testcg(self):
        b = self.b.copy()
        out = np.zeros(4, np.int32)
        nr = ctypes.c_double()
        lib().or_testcg(self._h, _p(b), _p(out), ctypes.byref(nr))
        return dict(count_pass=int(out[0]), count_fail=int(out[1]),
                    niters_max_no_prec=int(out[2]), niters_max_prec=int(out[3]),
                    normr=nr.value)

    def testsymmetry(self, xr, yr):
        out = np.zeros(5)
        lib().or_testsymmetry(self._h, _p(self.b), _p(self.xexact),
                              _p(np.ascontiguousarray(xr, np.float64)),
                              _p(np.ascontiguousarray(yr, np.float64)), _p(out))
        return dict(depsym_spmv=out[0], depsym_mg=out[1], count_fail=int(out[2]),
                    residuals=(out[3], out[4]))


# ---- mixed precision ---------------------------------------------------------
_MIXED_SIGS = {
    "or_half_round": (ctypes.c_float, [ctypes.c_double]),
    "or_symgs32": (None, [ctypes.c_int64] + [ctypes.c_void_p] * 7),
    "or_spmv32": (None, [ctypes.c_int64] + [ctypes.c_void_p] * 6),
    "or_mg32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "or_cg_mixed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
}


def _mixed_lib():
    L = lib()
    if not getattr(L, "_mixed_sigs", False):
        for name, (res, args) in _MIXED_SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._mixed_sigs = True
    return L


def half_round(v):
    return _mixed_lib().or_half_round(float(v))


def symgs32(m, r, x, order=None):
    r = np.ascontiguousarray(r, np.float32)
    x = np.array(x, np.float32, copy=True)
    o = None if order is None else np.ascontiguousarray(order, np.int32)
    _mixed_lib().or_symgs32(x.size, _p(m["nnz_row"]), _p(m["cols"]), _p(m["vals"]),
                            _p(m["diag_slot"]), _p(o), _p(r), _p(x))
    return x


def spmv32(m, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(m["nnz_row"].size, np.float32)
    _mixed_lib().or_spmv32(y.size, _p(m["nnz_row"]), _p(m["cols"]), _p(m["vals"]),
                           _p(m["diag_slot"]), _p(x), _p(y))
    return y


def _hier_mg32(self, r, level=0):
    r = np.ascontiguousarray(r, np.float32)
    x = np.zeros(self.nrow(level), np.float32)
    _mixed_lib().or_mg32(self._h, level, _p(r), _p(x))
    return x


def _hier_cg_mixed(self, b=None, x=None, max_iter=50, tol=0.0, recompute=10):
    b = self.b.copy() if b is None else np.ascontiguousarray(b, np.float64)
    x = np.zeros(self.nrow(0)) if x is None else np.array(x, np.float64, copy=True)
    traj = np.zeros(max_iter + 1)
    ni, nr, nr0 = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    _mixed_lib().or_cg_mixed(self._h, _p(b), _p(x), max_iter, tol, recompute, ctypes.byref(ni),
                             ctypes.byref(nr), ctypes.byref(nr0), _p(traj))
    return dict(x=x, niters=ni.value, normr=nr.value, normr0=nr0.value,
                traj=traj[: ni.value + 1])


Hierarchy.mg32 = _hier_mg32
Hierarchy.cg_mixed = _hier_cg_mixed
