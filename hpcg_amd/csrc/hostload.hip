// hostload.hip -- a level given in the reference's host storage (arrays of
// row pointers, SparseMatrix.hpp:39-77) turned into the natural-order ELL on
// the device.
//
// The rows travel in chunks: worker threads pack a chunk's rows into a
// pinned staging buffer (one memcpy when the rows lie back to back in one
// block, as the reference's generator allocates them; row by row otherwise),
// the buffer is copied to the device while the next chunk is packed into the
// other buffer, and a kernel writes the chunk's rows into the column-major
// ELL with the usual padding (column = the row, value 0).
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace hpcgx {
namespace {

constexpr int64_t kChunkRows = 1 << 21;  // rows per staging buffer (648 MB)

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// rows [r0, r0 + rows) from row-major staging ([row][27], the first
// nnz_row entries valid) into the column-major ELL of n rows
__global__ __launch_bounds__(kBlock) void k_rows_to_ell(int64_t rows, int64_t r0, int64_t n,
                                                        const int32_t *__restrict__ rc,
                                                        const double *__restrict__ rv,
                                                        const uint8_t *__restrict__ nnz_row,
                                                        int32_t *__restrict__ col,
                                                        double *__restrict__ val) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows) return;
  const int64_t i = r0 + t;
  const int m = nnz_row[i];
  for (int j = 0; j < kSlots; ++j) {
    const bool on = j < m;
    col[(int64_t)j * n + i] = on ? rc[t * kSlots + j] : (int32_t)i;
    val[(int64_t)j * n + i] = on ? rv[t * kSlots + j] : 0.0;
  }
}

// f(begin, end) over [0, n) split across the host's threads
template <class F>
void parallel_rows(int64_t n, F &&f) {
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < (1 << 16) || nt == 1) {
    f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (int k = 0; k < nt; ++k)
    pool.emplace_back([&, k] { f(n * k / nt, n * (k + 1) / nt); });
  for (auto &t : pool) t.join();
}

struct Pinned {
  void *p = nullptr;
  explicit Pinned(size_t bytes) { HPCGX_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault)); }
  ~Pinned() { (void)hipHostFree(p); }
};

}  // namespace

void upload_host_level(const hpcgx_host_level &H, NaturalLevel &N, hipStream_t s) {
  const int64_t n = H.nrow;
  // row lengths in [1, 27]; are the rows one block each for columns and values?
  std::atomic<bool> bad_len{false}, split{false};
  const int32_t *c0 = H.mtxIndL[0];
  const double *v0 = H.matrixValues[0];
  parallel_rows(n, [&](int64_t b, int64_t e) {
    bool bl = false, sp = false;
    for (int64_t i = b; i < e; ++i) {
      const int m = H.nonzerosInRow[i];
      bl |= m < 1 || m > kSlots;
      sp |= H.mtxIndL[i] != c0 + kSlots * i || H.matrixValues[i] != v0 + kSlots * i;
    }
    if (bl) bad_len = true;
    if (sp) split = true;
  });
  if (bad_len) throw Error(HPCGX_EUNSUP, "row with 0 or more than 27 entries");
  HPCGX_HIP(hipMemcpyAsync(N.nnz_row, H.nonzerosInRow, n, hipMemcpyHostToDevice, s));

  const int64_t CH = std::min<int64_t>(kChunkRows, n);
  const size_t cb = sizeof(int32_t) * kSlots * CH, vb = sizeof(double) * kSlots * CH;
  Pinned hc0(cb), hc1(cb), hv0(vb), hv1(vb);
  TempBufD dc0(cb), dc1(cb), dv0(vb), dv1(vb);
  int32_t *hc[2] = {static_cast<int32_t *>(hc0.p), static_cast<int32_t *>(hc1.p)};
  double *hv[2] = {static_cast<double *>(hv0.p), static_cast<double *>(hv1.p)};
  int32_t *dc[2] = {static_cast<int32_t *>(dc0.p), static_cast<int32_t *>(dc1.p)};
  double *dv[2] = {static_cast<double *>(dv0.p), static_cast<double *>(dv1.p)};
  hipEvent_t used[2];
  for (auto &e : used) HPCGX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  try {
    for (int64_t r0 = 0, k = 0; r0 < n; r0 += CH, ++k) {
      const int q = (int)(k & 1);
      const int64_t rows = std::min(CH, n - r0);
      if (k >= 2) HPCGX_HIP(hipEventSynchronize(used[q]));  // staging q free again
      parallel_rows(rows, [&](int64_t b, int64_t e) {
        if (!split) {
          std::memcpy(hc[q] + b * kSlots, c0 + (r0 + b) * kSlots, sizeof(int32_t) * kSlots * (e - b));
          std::memcpy(hv[q] + b * kSlots, v0 + (r0 + b) * kSlots, sizeof(double) * kSlots * (e - b));
          return;
        }
        for (int64_t t = b; t < e; ++t) {
          const int m = H.nonzerosInRow[r0 + t];
          std::memcpy(hc[q] + t * kSlots, H.mtxIndL[r0 + t], sizeof(int32_t) * m);
          std::memcpy(hv[q] + t * kSlots, H.matrixValues[r0 + t], sizeof(double) * m);
        }
      });
      HPCGX_HIP(hipMemcpyAsync(dc[q], hc[q], sizeof(int32_t) * kSlots * rows,
                               hipMemcpyHostToDevice, s));
      HPCGX_HIP(hipMemcpyAsync(dv[q], hv[q], sizeof(double) * kSlots * rows,
                               hipMemcpyHostToDevice, s));
      k_rows_to_ell<<<grid_for(rows), kBlock, 0, s>>>(rows, r0, n, dc[q], dv[q], N.nnz_row, N.col,
                                                     N.val);
      HPCGX_LAUNCHED();
      HPCGX_HIP(hipEventRecord(used[q], s));
    }
    HPCGX_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    for (auto &e : used) (void)hipEventDestroy(e);
    throw;
  }
  for (auto &e : used) (void)hipEventDestroy(e);
}

}  // namespace hpcgx
