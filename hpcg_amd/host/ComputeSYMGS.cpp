// ComputeSYMGS.cpp -- one symmetric Gauss-Seidel sweep on the GPU (replaces
// the reference's ComputeSYMGS.cpp:50-55).  The optimized ordering is the
// multicolour one; in the natural order the result equals
// ComputeSYMGS_ref.cpp:54-103 bit for bit.
#include "DeviceProblem.hpp"
#include "OptimizedKernels.hpp"

namespace hpcg_amd {

int DeviceSYMGS(const SparseMatrix &A, const Vector &r, Vector &x, int ordering) {
  assert(x.localLength == A.localNumberOfColumns);  // ComputeSYMGS_ref.cpp:56
  const DeviceLevel *L = LevelFor(A, "ComputeSYMGS");
  if (!L) return HPCGX_EINVAL;
  double *dr = L->ctx->vec[0], *dx = L->ctx->vec[1];
  int rc = Upload(*L, r, dr);
  if (!rc) rc = Upload(*L, x, dx);
  if (rc) return rc;
  rc = hpcgx_symgs(L->ctx->problem, L->level, ordering, dr, dx, nullptr);
  if (rc) return Report("ComputeSYMGS", rc);
  return Download(*L, dx, x);
}

}  // namespace hpcg_amd

int ComputeSYMGS(const SparseMatrix &A, const Vector &r, Vector &x) {
  return hpcg_amd::DeviceSYMGS(A, r, x, hpcg_amd::optimized_ordering);
}
