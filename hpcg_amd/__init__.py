"""hpcg_amd -- MI355X-native implementation of the HPCG 3.1 PCG iteration.

The compute path is libhpcg_amd.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/hpcg_amd.h).  This package is its host-side mirror of the
reference's optimisable interface (ComputeSPMV, ComputeSYMGS, ComputeMG,
ComputeDotProduct, ComputeWAXPBY, CG, OptimizeProblem) plus the reference's
flop/byte model.  There is no CPU fallback: the library must be built.
"""
from ._lib import (CG_GRAPH, CG_TIMERS, KEEP_SCHEDULE, ORDER_LEXICOGRAPHIC, ORDER_MULTICOLOR, HpcgxError,
                   exported_symbols, lib)
from .problem import Problem
from . import model
from .fleet import fleet_dims

__all__ = ["Problem", "HpcgxError", "ORDER_LEXICOGRAPHIC", "ORDER_MULTICOLOR", "model", "lib",
           "exported_symbols", "KEEP_SCHEDULE", "CG_TIMERS", "CG_GRAPH", "fleet_dims"]
