// CG.cpp -- preconditioned conjugate gradients on the GPU (replaces the
// reference's CG.cpp:59-143).  b and x cross to the device once per call;
// every iteration (V-cycle, SpMV, dots, updates) stays on the device, with
// the reductions fused into the kernels that produce their operands.
// times[0] covers the whole call, copies included; times[1..5] (dot,
// WAXPBY, SpMV, all-reduce = 0, MG) come from HIP events around each phase
// when cg_phase_timers is set (the default), read once per call.
#include <chrono>
#include <cstdio>

#include "DeviceProblem.hpp"
#include "OptimizedKernels.hpp"

namespace hpcg_amd {

int DeviceCG(const SparseMatrix &A, const Vector &b, Vector &x, int max_iter,
             double tolerance, int &niters, double &normr, double &normr0,
             double *times, bool doPreconditioning, int ordering) {
  const DeviceLevel *L = LevelFor(A, "CG");
  if (!L) return HPCGX_EINVAL;
  const auto t0 = std::chrono::steady_clock::now();
  if (L->level != 0) {
    std::fprintf(stderr, "hpcg_amd: CG: A must be the fine level\n");
    return HPCGX_EINVAL;
  }
  double *db = L->ctx->vec[0], *dx = L->ctx->vec[1];
  int rc = Upload(*L, b, db);
  if (!rc) rc = Upload(*L, x, dx);
  if (rc) return rc;
  double t[6] = {0, 0, 0, 0, 0, 0};
  // without phase timers, iterations 2.. replay a captured graph
  rc = hpcgx_cg(L->ctx->problem, ordering, db, dx, max_iter, tolerance,
                doPreconditioning ? 1 : 0,
                (cg_phase_timers ? HPCGX_CG_TIMERS : 0) | HPCGX_CG_GRAPH, &niters, &normr,
                &normr0, t, nullptr);
  if (rc) return Report("CG", rc);
  rc = Download(*L, dx, x);
  if (times) {
    times[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int k = 1; k < 6; ++k) times[k] += t[k];
  }
  return rc;
}

}  // namespace hpcg_amd

int CG(const SparseMatrix &A, CGData &data, const Vector &b, Vector &x, const int max_iter,
       const double tolerance, int &niters, double &normr, double &normr0, double *times,
       bool doPreconditioning) {
  (void)data;  // the device problem holds r, z, p and Ap
  return hpcg_amd::DeviceCG(A, b, x, max_iter, tolerance, niters, normr, normr0, times,
                            doPreconditioning, hpcg_amd::optimized_ordering);
}
