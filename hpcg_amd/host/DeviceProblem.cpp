// DeviceProblem.cpp -- device problem lifetime and host<->device copies for
// the C++ surface.
#include "DeviceProblem.hpp"

#include <algorithm>

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <vector>

namespace hpcg_amd {

int optimized_ordering = HPCGX_ORDER_MULTICOLOR;
bool cg_phase_timers = true;
bool grid_knowledge = true;

namespace {
DeviceContext *g_current = nullptr;

int hip_report(const char *where, hipError_t e) {
  std::fprintf(stderr, "hpcg_amd: %s: %s\n", where, hipGetErrorString(e));
  return HPCGX_EHIP;
}

void free_context(DeviceContext *c) {
  if (!c) return;
  if (c->stage) (void)hipFree(c->stage);
  for (double *v : c->vec)
    if (v) (void)hipFree(v);
  for (int k = 0; k < 2; ++k) {
    if (c->pinned[k]) (void)hipHostFree(c->pinned[k]);
    if (c->copied[k]) (void)hipEventDestroy(c->copied[k]);
  }
  if (c->problem) (void)hpcgx_problem_destroy(c->problem);
  if (g_current == c) g_current = nullptr;
  delete c;
}
}  // namespace

int Report(const char *where, int rc) {
  std::fprintf(stderr, "hpcg_amd: %s: %s\n", where, hpcgx_last_error());
  return rc ? rc : HPCGX_EHIP;
}

DeviceContext *EnsureDeviceProblem(SparseMatrix &A) {
  if (A.optimizationData) return static_cast<DeviceLevel *>(A.optimizationData)->ctx;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<hpcgx_host_level> levels;
  std::vector<SparseMatrix *> mats;
  for (SparseMatrix *M = &A; M; M = M->Ac) {
    if (levels.size() == 8) {
      std::fprintf(stderr, "hpcg_amd: OptimizeProblem: more than 8 levels\n");
      return nullptr;
    }
    hpcgx_host_level h{};
    if (grid_knowledge) {  // lets the setup propose the schedule from the grid
      h.nx = M->geom->nx;
      h.ny = M->geom->ny;
      h.nz = M->geom->nz;
    }
    h.nrow = M->localNumberOfRows;
    h.nonzerosInRow = M->nonzerosInRow;
    h.mtxIndL = M->mtxIndL;
    h.matrixValues = M->matrixValues;
    h.matrixDiagonal = M->matrixDiagonal;
    h.f2cOperator = (M->Ac && M->mgData) ? M->mgData->f2cOperator : nullptr;
    levels.push_back(h);
    mats.push_back(M);
    if (M->Ac && !M->mgData) break;  // a coarse level without transfer data ends the chain
  }
  auto *c = new DeviceContext{};
  int rc = hpcgx_optimize_problem(levels.data(), (int)levels.size(), 0, &c->problem);
  if (rc) {
    Report("OptimizeProblem", rc);
    free_context(c);
    return nullptr;
  }
  if (!grid_knowledge && (rc = hpcgx_tune(c->problem, "affine8", 0.0))) {
    Report("OptimizeProblem", rc);
    free_context(c);
    return nullptr;
  }
  c->nlevels = (int)levels.size();
  c->fine = &A;
  for (int l = 0; l < c->nlevels; ++l) {
    c->nrow[l] = levels[l].nrow;
    c->levels[l] = DeviceLevel{c, l};
  }
  const size_t bytes = sizeof(double) * (size_t)c->nrow[0];
  hipError_t e = hipMalloc(&c->stage, bytes);
  for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipMalloc(&c->vec[k], bytes);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    e = hipHostMalloc(reinterpret_cast<void **>(&c->pinned[k]), sizeof(double) * kCopyChunk,
                      hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->copied[k], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    hip_report("OptimizeProblem: scratch vectors", e);
    free_context(c);
    return nullptr;
  }
  for (int l = 0; l < c->nlevels; ++l) mats[l]->optimizationData = &c->levels[l];
  c->build_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  g_current = c;
  return c;
}

const DeviceLevel *LevelOf(const SparseMatrix &A) {
  return static_cast<const DeviceLevel *>(A.optimizationData);
}

const DeviceLevel *LevelFor(const SparseMatrix &A, const char *where) {
  if (const DeviceLevel *L = LevelOf(A)) return L;
  if (A.geom && A.localNumberOfRows > 0 && A.nonzerosInRow) {
    // the caller's matrix stays logically const: only optimizationData,
    // the slot the reference reserves for this, is written
    if (EnsureDeviceProblem(const_cast<SparseMatrix &>(A))) return LevelOf(A);
  }
  std::fprintf(stderr, "hpcg_amd: %s: matrix has no device problem\n", where);
  return nullptr;
}

DeviceContext *CurrentContext() { return g_current; }

int Upload(const DeviceLevel &L, const Vector &v, double *dev) {
  const int64_t n = L.ctx->nrow[L.level];
  if (v.localLength < n) {
    std::fprintf(stderr, "hpcg_amd: vector of %d values on a level of %lld rows\n",
                 v.localLength, (long long)n);
    return HPCGX_EINVAL;
  }
  DeviceContext &c = *L.ctx;
  hipError_t e = hipSuccess;
  for (int64_t off = 0, k = 0; off < n && e == hipSuccess; off += kCopyChunk, ++k) {
    const int q = (int)(k & 1);
    const int64_t m = std::min<int64_t>(kCopyChunk, n - off);
    // buffer q free again: its last copy (in this call or an earlier one) is done
    e = hipEventSynchronize(c.copied[q]);
    if (e != hipSuccess) break;
    double *dst = c.pinned[q];
    const double *src = v.values + off;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) dst[i] = src[i];
    e = hipMemcpyAsync(c.stage + off, dst, sizeof(double) * m, hipMemcpyHostToDevice, nullptr);
    if (e == hipSuccess) e = hipEventRecord(c.copied[q], nullptr);
  }
  if (e != hipSuccess) return hip_report("host to device copy", e);
  int rc = hpcgx_to_device_order(c.problem, L.level, c.stage, dev, nullptr);
  return rc ? Report("to_device_order", rc) : 0;
}

int Download(const DeviceLevel &L, const double *dev, Vector &v) {
  const int64_t n = L.ctx->nrow[L.level];
  if (v.localLength < n) {
    std::fprintf(stderr, "hpcg_amd: vector of %d values on a level of %lld rows\n",
                 v.localLength, (long long)n);
    return HPCGX_EINVAL;
  }
  DeviceContext &c = *L.ctx;
  int rc = hpcgx_to_natural_order(c.problem, L.level, dev, c.stage, nullptr);
  if (rc) return Report("to_natural_order", rc);
  // chunk k+1 crosses PCIe while the host unpacks chunk k
  const int64_t chunks = (n + kCopyChunk - 1) / kCopyChunk;
  auto issue = [&](int64_t k) {
    const int q = (int)(k & 1);
    const int64_t off = k * kCopyChunk, m = std::min<int64_t>(kCopyChunk, n - off);
    hipError_t r = hipMemcpyAsync(c.pinned[q], c.stage + off, sizeof(double) * m,
                                  hipMemcpyDeviceToHost, nullptr);
    return r == hipSuccess ? hipEventRecord(c.copied[q], nullptr) : r;
  };
  hipError_t e = chunks > 0 ? issue(0) : hipSuccess;
  for (int64_t k = 0; k < chunks && e == hipSuccess; ++k) {
    const int q = (int)(k & 1);
    e = hipEventSynchronize(c.copied[q]);
    if (e == hipSuccess && k + 1 < chunks) e = issue(k + 1);
    if (e != hipSuccess) break;
    const int64_t off = k * kCopyChunk, m = std::min<int64_t>(kCopyChunk, n - off);
    const double *src = c.pinned[q];
    double *dst = v.values + off;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) dst[i] = src[i];
  }
  return e == hipSuccess ? 0 : hip_report("device to host copy", e);
}

int UploadRaw(const double *host, double *dev, int64_t n) {
  hipError_t e = hipMemcpy(dev, host, sizeof(double) * n, hipMemcpyHostToDevice);
  return e == hipSuccess ? 0 : hip_report("host to device copy", e);
}

int DownloadRaw(const double *dev, double *host, int64_t n) {
  hipError_t e = hipMemcpy(host, dev, sizeof(double) * n, hipMemcpyDeviceToHost);
  return e == hipSuccess ? 0 : hip_report("device to host copy", e);
}

}  // namespace hpcg_amd

using namespace hpcg_amd;

void MatrixReplaceDiagonalOnDevice(SparseMatrix &A, const Vector &diagonal) {
  const DeviceLevel *L = LevelOf(A);
  if (!L) return;  // no device copy yet: OptimizeProblem will read the host values
  double *d = L->ctx->vec[3];
  int rc = Upload(*L, diagonal, d);
  if (!rc) rc = hpcgx_replace_diagonal(L->ctx->problem, L->level, d, nullptr);
  if (rc) {
    Report("ReplaceMatrixDiagonal", rc);
    std::abort();  // the device matrix would silently differ from the host one
  }
}

void MatrixReleaseDevice(SparseMatrix &A) {
  const DeviceLevel *L = LevelOf(A);
  if (!L) return;
  DeviceContext *c = L->ctx;
  if (c->fine != &A) return;  // released with the fine level
  for (SparseMatrix *M = &A; M; M = M->Ac) M->optimizationData = nullptr;
  (void)hipDeviceSynchronize();
  free_context(c);
}
