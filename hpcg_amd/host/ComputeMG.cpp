// ComputeMG.cpp -- the multigrid V-cycle on the GPU (replaces the
// reference's ComputeMG.cpp:33-38; same cycle as ComputeMG_ref.cpp:39-63:
// one pre- and one post-smoothing sweep, residual restriction by injection,
// prolongation, SYMGS alone on the coarsest level).  The whole cycle runs on
// the device problem's own coarse vectors.
#include "DeviceProblem.hpp"
#include "OptimizedKernels.hpp"

namespace hpcg_amd {

int DeviceMG(const SparseMatrix &A, const Vector &r, Vector &x, int ordering) {
  const DeviceLevel *L = LevelFor(A, "ComputeMG");
  if (!L) return HPCGX_EINVAL;
  double *dr = L->ctx->vec[0], *dx = L->ctx->vec[1];
  int rc = Upload(*L, r, dr);
  if (rc) return rc;
  rc = hpcgx_mg(L->ctx->problem, L->level, ordering, dr, dx, nullptr);
  if (rc) return Report("ComputeMG", rc);
  return Download(*L, dx, x);
}

}  // namespace hpcg_amd

int ComputeMG(const SparseMatrix &A, const Vector &r, Vector &x) {
  return hpcg_amd::DeviceMG(A, r, x, hpcg_amd::optimized_ordering);
}
