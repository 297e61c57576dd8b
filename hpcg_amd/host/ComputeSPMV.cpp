// ComputeSPMV.cpp -- y = A x on the GPU (replaces the reference's
// ComputeSPMV.cpp:40-45; the result equals ComputeSPMV_ref.cpp:47-72 bit for
// bit: same per-row summation order, no fused multiply-add).
#include "DeviceProblem.hpp"
#include "OptimizedKernels.hpp"

using namespace hpcg_amd;

int ComputeSPMV(const SparseMatrix &A, Vector &x, Vector &y) {
  assert(x.localLength >= A.localNumberOfColumns);
  assert(y.localLength >= A.localNumberOfRows);
  const DeviceLevel *L = LevelFor(A, "ComputeSPMV");
  if (!L) return HPCGX_EINVAL;
  double *dx = L->ctx->vec[0], *dy = L->ctx->vec[1];
  int rc = Upload(*L, x, dx);
  if (rc) return rc;
  rc = hpcgx_spmv(L->ctx->problem, L->level, dx, dy, nullptr);
  if (rc) return Report("ComputeSPMV", rc);
  return Download(*L, dy, y);
}
