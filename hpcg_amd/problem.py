"""Host mirror of the reference's optimisable PCG surface over the C ABI.

A :class:`Problem` owns a device-resident MG hierarchy (libhpcg_amd's
``hpcgx_problem``).  Its methods mirror the reference functions they replace
(ComputeSPMV, ComputeSYMGS, ComputeMG, ComputeDotProduct, ComputeWAXPBY, CG,
the restriction/prolongation pair, Replace/CopyMatrixDiagonal) and take
device vectors -- torch tensors, used only as device memory and streams --
in the problem's device row order.  ``to_device_order`` / ``to_natural_order``
convert at the host touch points.  Like the reference they return 0-or-raise:
a nonzero code from the library raises :class:`HpcgxError`.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (CG_GRAPH, CG_TIMERS, KEEP_SCHEDULE, ORDER_LEXICOGRAPHIC, ORDER_MULTICOLOR,
                   HostLevel, LevelInfo, call)

ORDERS = {"lexicographic": ORDER_LEXICOGRAPHIC, "level": ORDER_LEXICOGRAPHIC,
          "multicolor": ORDER_MULTICOLOR, "color": ORDER_MULTICOLOR,
          ORDER_LEXICOGRAPHIC: ORDER_LEXICOGRAPHIC, ORDER_MULTICOLOR: ORDER_MULTICOLOR}


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t, n=None, what="vector", dtype="float64"):
    """Device pointer of a contiguous CUDA tensor (float64 unless stated)
    of length >= n."""
    if t is None:
        return None
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("%s must be a CUDA tensor" % what)
    if t.dtype != getattr(torch, dtype) or not t.is_contiguous():
        raise TypeError("%s must be contiguous %s" % (what, dtype))
    if n is not None and t.numel() < n:
        raise ValueError("%s has %d entries, needs %d" % (what, t.numel(), n))
    return ctypes.c_void_p(t.data_ptr())


def _order(o):
    try:
        return ORDERS[o]
    except KeyError:
        raise ValueError("unknown ordering %r" % (o,))


def _flags(keep_schedule, band_lines, sweep_order=0):
    if not 0 <= band_lines <= 255:
        raise ValueError("band_lines must be in [0, 255]")
    if not 0 <= sweep_order <= 15:
        raise ValueError("sweep_order must be in [0, 15]")
    return (KEEP_SCHEDULE if keep_schedule else 0) | (band_lines << 8) | (sweep_order << 20)


class Problem:
    """Device-resident HPCG problem (all levels of the MG hierarchy)."""

    def __init__(self, handle):
        self._h = handle
        self.nlevels = _lib.lib().hpcgx_problem_nlevels(handle)
        self.levels = [self.info(l) for l in range(self.nlevels)]

    # ---- construction --------------------------------------------------
    @classmethod
    def generate(cls, nx, ny, nz, nlevels=4, keep_schedule=False, band_lines=0,
                 sweep_order=0):
        """GenerateProblem + GenerateCoarseProblem + OptimizeProblem on the
        device (main.cpp:133-144, OptimizeProblem.cpp:37).  band_lines: rows
        inside a colour run in bands of this many grid lines (0 = default,
        1 = natural order; see HPCGX_BAND_LINES).  sweep_order: visiting
        order of the colours (HPCGX_SWEEP_ORDER; 0 = default, 1 = id order)."""
        h = ctypes.c_void_p()
        call("hpcgx_generate_problem", nx, ny, nz, nlevels,
             _flags(keep_schedule, band_lines, sweep_order), ctypes.byref(h))
        return cls(h)

    @classmethod
    def from_reference_levels(cls, levels, keep_schedule=False, band_lines=0,
                              reversed_storage=False, sweep_order=0):
        """OptimizeProblem from host matrices in the reference's layout.

        ``levels`` is a list of dicts with numpy arrays ``nnz_row`` (n),
        ``cols`` (n*27, row i in [27i, 27i+27)), ``vals`` (n*27),
        ``diag_slot`` (n), optional ``f2c`` and dims ``nx, ny, nz``.  The
        arrays-of-row-pointers the reference's SparseMatrix holds are built
        over them.  reversed_storage: keep row i's entries in block n-1-i, so
        the row pointers are not one block in row order (as with rows
        allocated one by one).
        """
        keep = []
        arr = (HostLevel * len(levels))()
        for l, m in enumerate(levels):
            n = m["nnz_row"].size
            nnz = np.ascontiguousarray(m["nnz_row"], dtype=np.int8)
            cols = np.ascontiguousarray(m["cols"], dtype=np.int32)
            vals = np.ascontiguousarray(m["vals"], dtype=np.float64)
            rows = np.arange(n, dtype=np.int64)
            if reversed_storage:
                cols = np.ascontiguousarray(cols.reshape(n, 27)[::-1]).reshape(-1)
                vals = np.ascontiguousarray(vals.reshape(n, 27)[::-1]).reshape(-1)
                rows = rows[::-1].copy()
            ind_ptr = (cols.ctypes.data + 4 * 27 * rows).astype(np.uint64)
            val_ptr = (vals.ctypes.data + 8 * 27 * rows).astype(np.uint64)
            dia_ptr = (vals.ctypes.data + 8 * (27 * rows + m["diag_slot"].astype(np.int64))
                       ).astype(np.uint64)
            f2c = None if m.get("f2c") is None else np.ascontiguousarray(m["f2c"], np.int32)
            keep += [nnz, cols, vals, ind_ptr, val_ptr, dia_ptr, f2c]
            arr[l].nx, arr[l].ny, arr[l].nz = m.get("nx", 0), m.get("ny", 0), m.get("nz", 0)
            arr[l].nrow = n
            arr[l].nonzerosInRow = nnz.ctypes.data
            arr[l].mtxIndL = ind_ptr.ctypes.data
            arr[l].matrixValues = val_ptr.ctypes.data
            arr[l].matrixDiagonal = dia_ptr.ctypes.data
            arr[l].f2cOperator = None if f2c is None else f2c.ctypes.data
        h = ctypes.c_void_p()
        call("hpcgx_optimize_problem", arr, len(levels),
             _flags(keep_schedule, band_lines, sweep_order), ctypes.byref(h))
        del keep
        return cls(h)

    @classmethod
    def generate_fleet(cls, dims, nlevels=4, keep_schedule=False):
        """hpcgx_generate_fleet: independent problems on the grids in `dims`
        ([(nx, ny, nz)]) as one handle.  Level l is the block-diagonal union
        of the problems' level l (natural rows problem by problem); every
        level method treats it as one matrix, cg_fleet solves each problem
        with its own scalars."""
        q = len(dims)
        nx = np.ascontiguousarray([d[0] for d in dims], np.int32)
        ny = np.ascontiguousarray([d[1] for d in dims], np.int32)
        nz = np.ascontiguousarray([d[2] for d in dims], np.int32)
        h = ctypes.c_void_p()
        call("hpcgx_generate_fleet", q, nx.ctypes.data, ny.ctypes.data, nz.ctypes.data,
             nlevels, _flags(keep_schedule, 0), ctypes.byref(h))
        return cls(h)

    def check_problem(self, b=None, x=None, xexact=None):
        """CheckProblem (CheckProblem.cpp:50-151): every level's matrix and
        f2c, and b, x, xexact where given (device order), against the
        generated 27-point problem of the level's grids.  Raises HpcgxError
        (code 5, HPCGX_ECHECK) naming the first differing natural row."""
        n = self.nrow(0)
        return call("hpcgx_check_problem", self._h, _ptr(b, n, "b"), _ptr(x, n, "x"),
                    _ptr(xexact, n, "xexact"), _stream())

    def fleet_offsets(self, level=0):
        """Natural row offsets (count + 1) of the fleet's problems on `level`,
        or None for a single problem."""
        c = ctypes.c_int()
        call("hpcgx_fleet_info", self._h, level, ctypes.byref(c), None)
        if c.value == 0:
            return None
        off = np.zeros(c.value + 1, np.int64)
        call("hpcgx_fleet_info", self._h, level, ctypes.byref(c), off.ctypes.data)
        return off

    def cg_fleet(self, b, x, max_iter=50, tolerance=0.0, precondition=True,
                 ordering="multicolor", graph=False):
        """hpcgx_cg_fleet: CG on every problem of the fleet, each with its own
        scalars and stopping test.  Returns per-problem niters, normr, normr0
        (lists) and the times[] slots."""
        n = self.nrow(0)
        q = len(self.fleet_offsets()) - 1
        ni = (ctypes.c_int * q)()
        nr, nr0 = (ctypes.c_double * q)(), (ctypes.c_double * q)()
        t = (ctypes.c_double * 6)()
        call("hpcgx_cg_fleet", self._h, _order(ordering), _ptr(b, n, "b"), _ptr(x, n, "x"),
             int(max_iter), float(tolerance), int(bool(precondition)),
             CG_GRAPH if graph else 0, ni, nr, nr0, t, _stream())
        return dict(niters=list(ni), normr=list(nr), normr0=list(nr0), times=list(t))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hpcgx_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- queries -------------------------------------------------------
    def info(self, level=0):
        i = LevelInfo()
        call("hpcgx_problem_info", self._h, level, ctypes.byref(i))
        return dict(nx=i.nx, ny=i.ny, nz=i.nz, nrow=i.nrow, nnz=i.nnz, ncolors=i.ncolors,
                    nlevels_sched=i.nlevels_sched, ncoarse=i.ncoarse, wave=bool(i.wave),
                    nslabs=i.nslabs, col16=bool(i.col16), wide_tiles=i.wide_tiles,
                    byte_tiles=i.byte_tiles)

    def nrow(self, level=0):
        return self.levels[level]["nrow"]

    def memory_use(self):
        """OptimizeProblemMemoryUse: bytes of device memory held."""
        return _lib.lib().hpcgx_optimize_problem_memory_use(self._h)

    def schedule(self, level=0, which="level"):
        """Per-row level ids (which='level') or colours (which='color') in
        natural row order (needs keep_schedule=True)."""
        out = np.zeros(self.nrow(level), np.int32)
        call("hpcgx_get_schedule", self._h, level, {"level": 0, "color": 1}[which],
             out.ctypes.data_as(ctypes.c_void_p))
        return out

    # ---- vectors -------------------------------------------------------
    def empty(self, level=0):
        torch = _torch()
        return torch.empty(self.nrow(level), dtype=torch.float64, device="cuda")

    def zeros(self, level=0):
        torch = _torch()
        return torch.zeros(self.nrow(level), dtype=torch.float64, device="cuda")

    def vectors(self):
        """b, x, xexact of GenerateProblem_ref.cpp:178-180 in device order."""
        b, x, xe = self.empty(), self.empty(), self.empty()
        call("hpcgx_problem_vectors", self._h, _ptr(b), _ptr(x), _ptr(xe), _stream())
        return b, x, xe

    def to_device_order(self, natural, level=0):
        """natural-order vector (tensor or numpy) -> new device-order tensor."""
        torch = _torch()
        if not isinstance(natural, torch.Tensor):
            natural = torch.as_tensor(np.ascontiguousarray(natural, np.float64)).cuda()
        out = self.empty(level)
        call("hpcgx_to_device_order", self._h, level, _ptr(natural, self.nrow(level)),
             _ptr(out), _stream())
        return out

    def to_natural_order(self, devorder, level=0):
        out = self.empty(level)
        call("hpcgx_to_natural_order", self._h, level, _ptr(devorder, self.nrow(level)),
             _ptr(out), _stream())
        return out

    def to_numpy(self, devorder, level=0):
        """device-order tensor -> natural-order numpy array (synchronous)."""
        return self.to_natural_order(devorder, level).cpu().numpy()

    # ---- the optimisable surface ----------------------------------------
    def spmv(self, x, y, level=0):
        """ComputeSPMV (ComputeSPMV.cpp:40): y = A x."""
        n = self.nrow(level)
        return call("hpcgx_spmv", self._h, level, _ptr(x, n, "x"), _ptr(y, n, "y"), _stream())

    def symgs(self, r, x, level=0, ordering="lexicographic"):
        """ComputeSYMGS (ComputeSYMGS.cpp:50): one symmetric sweep on x."""
        n = self.nrow(level)
        return call("hpcgx_symgs", self._h, level, _order(ordering), _ptr(r, n, "r"),
                    _ptr(x, n, "x"), _stream())

    def mg(self, r, x, level=0, ordering="lexicographic"):
        """ComputeMG (ComputeMG.cpp:33): x = V-cycle(r)."""
        n = self.nrow(level)
        return call("hpcgx_mg", self._h, level, _order(ordering), _ptr(r, n, "r"),
                    _ptr(x, n, "x"), _stream())

    def restriction(self, rf, Axf, rc, level=0):
        """ComputeRestriction_ref (ComputeRestriction_ref.cpp:40)."""
        n, nc = self.nrow(level), self.nrow(level + 1)
        return call("hpcgx_restriction", self._h, level, _ptr(rf, n), _ptr(Axf, n),
                    _ptr(rc, nc), _stream())

    def residual_restriction(self, rf, xf, rc, level=0):
        """ComputeSPMV_ref + ComputeRestriction_ref fused (ComputeMG_ref.cpp:49-51)."""
        n, nc = self.nrow(level), self.nrow(level + 1)
        return call("hpcgx_residual_restriction", self._h, level, _ptr(rf, n), _ptr(xf, n),
                    _ptr(rc, nc), _stream())

    def prolongation(self, xc, xf, level=0):
        """ComputeProlongation_ref (ComputeProlongation_ref.cpp:38)."""
        n, nc = self.nrow(level), self.nrow(level + 1)
        return call("hpcgx_prolongation", self._h, level, _ptr(xc, nc), _ptr(xf, n), _stream())

    def dot(self, x, y, n=None):
        """ComputeDotProduct (ComputeDotProduct.cpp:41): returns the value."""
        n = x.numel() if n is None else n
        out = ctypes.c_double()
        call("hpcgx_dot_host", self._h, n, _ptr(x, n), _ptr(y, n), ctypes.byref(out), _stream())
        return out.value

    def dot_device(self, x, y, out, n=None):
        n = x.numel() if n is None else n
        return call("hpcgx_dot", self._h, n, _ptr(x, n), _ptr(y, n), _ptr(out, 1), _stream())

    @staticmethod
    def waxpby(alpha, x, beta, y, w, n=None):
        """ComputeWAXPBY (ComputeWAXPBY.cpp:42): w = alpha x + beta y."""
        n = x.numel() if n is None else n
        return call("hpcgx_waxpby", n, float(alpha), _ptr(x, n), float(beta), _ptr(y, n),
                    _ptr(w, n), _stream())

    def replace_diagonal(self, diag, level=0):
        """ReplaceMatrixDiagonal (SparseMatrix.hpp:139)."""
        return call("hpcgx_replace_diagonal", self._h, level, _ptr(diag, self.nrow(level)),
                    _stream())

    def copy_diagonal(self, level=0):
        """CopyMatrixDiagonal (SparseMatrix.hpp:126): new device-order tensor."""
        out = self.empty(level)
        call("hpcgx_copy_diagonal", self._h, level, _ptr(out), _stream())
        return out

    # ---- mixed precision (FP32 V-cycle on FP16 values) --------------------
    def prepare_mixed(self):
        return call("hpcgx_prepare_mixed", self._h)

    def zeros32(self, level=0):
        torch = _torch()
        return torch.zeros(self.nrow(level), dtype=torch.float32, device="cuda")

    def spmv32(self, x, y, level=0):
        n = self.nrow(level)
        return call("hpcgx_spmv32", self._h, level, _ptr(x, n, "x", "float32"),
                    _ptr(y, n, "y", "float32"), _stream())

    def symgs32(self, r, x, level=0, ordering="multicolor"):
        n = self.nrow(level)
        return call("hpcgx_symgs32", self._h, level, _order(ordering),
                    _ptr(r, n, "r", "float32"), _ptr(x, n, "x", "float32"), _stream())

    def mg32(self, r, x, level=0, ordering="multicolor"):
        n = self.nrow(level)
        return call("hpcgx_mg32", self._h, level, _order(ordering),
                    _ptr(r, n, "r", "float32"), _ptr(x, n, "x", "float32"), _stream())

    # ---- block PCG over s right-hand sides ([n][s] multivectors) ----------
    def to_device_order_multi(self, natural, level=0, dtype="float64"):
        """natural-order [n][s] array -> device-order [n][s] tensor."""
        torch = _torch()
        nat = np.asarray(natural, np.float64)
        cols = [self.to_device_order(nat[:, k], level) for k in range(nat.shape[1])]
        return torch.stack(cols, dim=1).to(getattr(torch, dtype)).contiguous()

    def to_numpy_multi(self, dev, level=0):
        """device-order [n][s] tensor -> natural-order [n][s] numpy array."""
        d = dev.double()
        return np.stack([self.to_numpy(d[:, k].contiguous(), level)
                         for k in range(d.shape[1])], axis=1)

    def spmv_block(self, X, Y, level=0):
        n, s = X.shape
        return call("hpcgx_spmv_block", self._h, level, s, _ptr(X, n * s, "X"),
                    _ptr(Y, n * s, "Y"), _stream())

    def mg32_block(self, R, X, level=0, ordering="multicolor"):
        n, s = R.shape
        return call("hpcgx_mg32_block", self._h, level, _order(ordering), s,
                    _ptr(R, n * s, "R", "float32"), _ptr(X, n * s, "X", "float32"), _stream())

    def cg_block_mixed(self, B, X, max_iter=50, tolerance=0.0, recompute=10,
                       ordering="multicolor"):
        """hpcgx_cg_block_mixed: s systems at once (B, X are [n][s], device
        order).  Returns niters and per-system normr, normr0."""
        n, s = B.shape
        ni = ctypes.c_int()
        nr, nr0 = (ctypes.c_double * s)(), (ctypes.c_double * s)()
        t = (ctypes.c_double * 6)()
        call("hpcgx_cg_block_mixed", self._h, _order(ordering), s, _ptr(B, n * s, "B"),
             _ptr(X, n * s, "X"), int(max_iter), float(tolerance), int(recompute),
             ctypes.byref(ni), nr, nr0, t, _stream())
        return dict(niters=ni.value, normr=list(nr), normr0=list(nr0), times=list(t))

    def cg_mixed(self, b, x, max_iter=50, tolerance=0.0, recompute=10, ordering="multicolor"):
        """FP64 PCG with the FP32 V-cycle (FP16 values) as preconditioner and
        r = b - A x recomputed in FP64 every `recompute` iterations."""
        n = self.nrow(0)
        ni, nr, nr0 = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        t = (ctypes.c_double * 6)()
        call("hpcgx_cg_mixed", self._h, _order(ordering), _ptr(b, n, "b"), _ptr(x, n, "x"),
             int(max_iter), float(tolerance), int(recompute), ctypes.byref(ni),
             ctypes.byref(nr), ctypes.byref(nr0), t, _stream())
        return dict(niters=ni.value, normr=nr.value, normr0=nr0.value, times=list(t))

    def tune(self, key, value):
        """Set a tuning knob ("col16", "col8", "spmv_dot_blocks", "spmv_order", "fleet_blocks",
        "wave", "wave_min_rows",
        "wave_slab_factor", "wave_rows", "xcd_remap", "pair_rows",
        "stream_min_bytes"); see include/hpcg_amd.h."""
        return call("hpcgx_tune", self._h, key.encode(), float(value))

    def autotune(self, ordering="multicolor", iterations=3):
        """hpcgx_autotune: time the candidate knob values and keep the
        fastest; returns the report lines."""
        buf = ctypes.create_string_buffer(4096)
        call("hpcgx_autotune", self._h, _order(ordering), int(iterations), buf, len(buf))
        return buf.value.decode().splitlines()

    def check(self):
        """Wait for the stream; raise if a wavefront sweep gave up waiting."""
        return call("hpcgx_check_status", self._h, _stream())

    def kernel_timing(self, enable=True):
        """Start (clearing the log) or stop the event log of the finest
        level's multicolour Gauss-Seidel launches."""
        return call("hpcgx_kernel_timing", self._h, int(bool(enable)))

    def kernel_timing_read(self):
        """(total milliseconds, launches) of the logged launches."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        call("hpcgx_kernel_timing_read", self._h, ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def cg(self, b, x, max_iter=50, tolerance=0.0, precondition=True,
           ordering="multicolor", timers=False, times=None, graph=False):
        """CG (CG.cpp:59-143).  x is updated in place; returns niters, normr,
        normr0 and the times[] slots accumulated by the call.  graph=True
        replays a captured hipGraph of the iteration (HPCGX_CG_GRAPH)."""
        n = self.nrow(0)
        ni, nr, nr0 = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        t = (ctypes.c_double * 6)(*(times if times is not None else [0.0] * 6))
        call("hpcgx_cg", self._h, _order(ordering), _ptr(b, n, "b"), _ptr(x, n, "x"),
             int(max_iter), float(tolerance), int(bool(precondition)),
             (CG_TIMERS if timers else 0) | (CG_GRAPH if graph else 0),
             ctypes.byref(ni), ctypes.byref(nr),
             ctypes.byref(nr0), t, _stream())
        return dict(niters=ni.value, normr=nr.value, normr0=nr0.value, times=list(t))
