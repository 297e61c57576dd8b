// ComputeWAXPBY.cpp -- w = alpha x + beta y on the GPU (replaces the
// reference's ComputeWAXPBY.cpp:42-48; the alpha == 1 and beta == 1 forms of
// ComputeWAXPBY_ref.cpp:52-67 give the same bits).  w may alias x or y.
#include <cstdio>

#include "DeviceProblem.hpp"
#include "OptimizedKernels.hpp"

using namespace hpcg_amd;

int ComputeWAXPBY(const local_int_t n, const double alpha, const Vector &x,
                  const double beta, const Vector &y, Vector &w, bool &isOptimized) {
  (void)isOptimized;
  assert(x.localLength >= n && y.localLength >= n && w.localLength >= n);
  DeviceContext *c = CurrentContext();
  if (!c) {
    std::fprintf(stderr, "hpcg_amd: ComputeWAXPBY: called before OptimizeProblem\n");
    return HPCGX_EINVAL;
  }
  if (n < 0 || n > c->nrow[0]) {
    std::fprintf(stderr, "hpcg_amd: ComputeWAXPBY: n out of range\n");
    return HPCGX_EINVAL;
  }
  double *dx = c->vec[0], *dy = c->vec[1], *dw = c->vec[2];
  int rc = UploadRaw(x.values, dx, n);
  if (rc) return rc;
  if (y.values == x.values) {
    dy = dx;
  } else if ((rc = UploadRaw(y.values, dy, n))) {
    return rc;
  }
  rc = hpcgx_waxpby(n, alpha, dx, beta, dy, dw, nullptr);
  if (rc) return Report("ComputeWAXPBY", rc);
  return DownloadRaw(dw, w.values, n);
}
