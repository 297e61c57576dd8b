// kernels.hip -- HIP kernels of the FP64 PCG path for gfx950.
//
// Every row kernel keeps the reference's per-row arithmetic: terms summed in
// slot order j = 0..26 with separate multiply and add (the library is built
// with -ffp-contract=off) and a correctly rounded division, so SpMV, the
// Gauss-Seidel sweeps, restriction and WAXPBY reproduce the reference bit for
// bit.  Dot products use a fixed-shape two-level tree, so they are
// bit-identical from call to call (TestNorms variance 0) and differ from the
// reference's sequential sum only by reassociation.
#include <algorithm>

#include "internal.hpp"

namespace hpcgx {

namespace {

inline unsigned grid_for(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}

__device__ inline int64_t gtid() {
  return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}

// ---------------------------------------------------------------------------
// BLAS-1 and data movement
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_fill(double *x, int64_t n, double v) {
  const int64_t i = gtid();
  if (i < n) x[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_copy(const double *__restrict__ x,
                                                  double *__restrict__ y,
                                                  int64_t n) {
  const int64_t i = gtid();
  if (i < n) y[i] = x[i];
}

__global__ __launch_bounds__(kBlock) void k_gather(
    const int32_t *__restrict__ idx, const double *__restrict__ src,
    double *__restrict__ dst, int64_t n) {
  const int64_t i = gtid();
  if (i < n) dst[i] = src[idx[i]];
}

// ComputeWAXPBY_ref.cpp:52-67: the three forms round differently, so the
// form is a template parameter chosen on the host exactly as the reference
// chooses it.  w may alias x or y (each element is read before it is written
// by the same thread).
// Workgroup b takes elements [b * kWaxE * kBlock, ...), thread q the
// elements b * kWaxE * kBlock + q + e * kBlock: every load of a thread is
// issued before its stores.  STREAM (vectors far larger than the last-level
// cache): non-temporal loads and stores.
constexpr int kWaxE = 4;
template <int FORM, bool STREAM>
__global__ __launch_bounds__(kBlock) void k_waxpby(int64_t n, double alpha,
                                                    const double *x,
                                                    double beta,
                                                    const double *y,
                                                    double *w) {
  const int64_t t0 = (int64_t)blockIdx.x * (kBlock * kWaxE) + threadIdx.x;
  double xv[kWaxE], yv[kWaxE];
#pragma unroll
  for (int e = 0; e < kWaxE; ++e) {
    const int64_t i = t0 + (int64_t)e * kBlock;
    xv[e] = i >= n ? 0.0 : STREAM ? __builtin_nontemporal_load(x + i) : x[i];
    yv[e] = i >= n ? 0.0 : STREAM ? __builtin_nontemporal_load(y + i) : y[i];
  }
#pragma unroll
  for (int e = 0; e < kWaxE; ++e) {
    const int64_t i = t0 + (int64_t)e * kBlock;
    if (i >= n) break;
    double v;
    if (FORM == 0)
      v = xv[e] + beta * yv[e];
    else if (FORM == 1)
      v = alpha * xv[e] + yv[e];
    else
      v = alpha * xv[e] + beta * yv[e];
    if (STREAM)
      __builtin_nontemporal_store(v, w + i);
    else
      w[i] = v;
  }
}

// ---------------------------------------------------------------------------
// Sparse row kernels on the column-major 27-slot ELL
// ---------------------------------------------------------------------------
// Matrix loads.  STREAM = the level's matrix is far larger than the 256 MB
// last-level cache, so it is read with non-temporal loads that do not evict
// the x vector; small levels read it through the caches, where it can stay
// resident between the sweeps of a smoother step.
template <bool STREAM, class T>
__device__ inline T mload(const T *p) {
  if (STREAM) return __builtin_nontemporal_load(p);
  return *p;
}
// A result the kernel never reads back: streamed past L2 on large levels so
// it does not evict the gathered vector.  The CG vector kernels stream their
// vectors from kStreamRows rows on (16 Mi rows: 128 MiB per FP64 vector).
template <bool STREAM, class T>
__device__ inline void mstore(T *p, T v) {
  if (STREAM)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// The hardware hands workgroup b to XCD b % 8.  Remapped, XCD x walks its own
// contiguous eighth of the logical blocks, so each XCD's L2 works on one
// narrow window of the gathered vector instead of all eight sharing it.
__device__ inline int64_t xcd_block(int64_t b, int64_t G, bool on) {
  if (!on) return b;
  const int64_t G8 = G - G % 8;
  if (b >= G8) return b;
  return (b % 8) * (G8 / 8) + b / 8;
}

// Sum over the workgroup in a fixed tree: 64-lane shuffles, then the four
// wavefront sums in order.  Result valid in thread 0.
__device__ inline double block_sum(double v) {
  __shared__ double wsum[kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wsum[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
  }
  __syncthreads();
  return t;
}

// y_i = sum_j a_ij x_col(i,j)  (ComputeSPMV_ref.cpp:62-69)
template <bool STREAM>
__device__ inline double row_dot(const int32_t *__restrict__ col,
                                 const double *__restrict__ val,
                                 const double *__restrict__ x, int64_t n,
                                 int64_t i) {
  double sum = 0.0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t k = (int64_t)j * n + i;
    sum += mload<STREAM>(val + k) * x[mload<STREAM>(col + k)];
  }
  return sum;
}

// SpMV over row chunks (<= kBlock rows each).  The chunk list interleaves the
// colour blocks (chunk k of every colour before chunk k+1 of any), so the x
// values gathered from all colour blocks stay within a narrow window.
template <bool STREAM, bool PAIRS = false>
__device__ inline double row_dot16(const Col16 &C, const int32_t *__restrict__ col,
                                   const double *__restrict__ val,
                                   const double *__restrict__ x, int64_t n,
                                   int64_t i, int64_t end = 0);

// C16: 16-bit column ids (chunks start at multiples of kTile).
template <bool STREAM, bool C16>
__global__ __launch_bounds__(kBlock) void k_spmv(
    const int64_t *__restrict__ chunks, int64_t n, Col16 C,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ x, double *__restrict__ y, bool xcd) {
  const int64_t b = xcd_block(blockIdx.x, gridDim.x, xcd);
  const int64_t i = chunks[2 * b] + threadIdx.x;
  if (i < chunks[2 * b + 1])
    y[i] = C16 ? row_dot16<STREAM>(C, col, val, x, n, i) : row_dot<STREAM>(col, val, x, n, i);
}

// One Gauss-Seidel row update in the reference's form
// (ComputeSYMGS_ref.cpp:71-80): subtract every stored term including the
// diagonal one, add the diagonal term back, divide by the diagonal.  The
// diagonal value is taken from the row's own slots: the stored entry whose
// column is the row (unused slots hold the value 0, and a stored zero
// diagonal would make the reference divide by zero as well), so no separate
// diagonal array is read.
template <bool STREAM>
__device__ inline double gs_row(int64_t i, int64_t n,
                                const int32_t *__restrict__ col,
                                const double *__restrict__ val,
                                const double *__restrict__ r, double *x) {
  double sum = r[i];
  double d = 0.0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t k = (int64_t)j * n + i;
    const double v = mload<STREAM>(val + k);
    const int32_t c = mload<STREAM>(col + k);
    sum -= v * x[c];
    if (c == i && v != 0.0) d = v;
  }
  sum += x[i] * d;
  x[i] = sum / d;
  return sum / d;
}

// Narrow column ids (Level::col16 / off8).  The caller's wavefront lies in
// one tile, so the tile id, its class and its slot bases are uniform (read
// through the scalar cache).  narrow_row() calls f(j, k, c) for every slot j
// in order, with k = j*n + i and c the slot's column.
__device__ inline int tile_of(int64_t i) {
  return __builtin_amdgcn_readfirstlane((int)(i / kTile));
}

template <bool STREAM, class F>
__device__ inline void narrow_row(const Col16 &C, int t, int info, int64_t n, int64_t i,
                                  F &&f) {
  const int32_t *__restrict__ base = C.base + t;
  if (info == kTile8 && C.use8 && C.amask && C.amask[t] == (1u << kSlotWords) - 1u) {
    // every offset word of the tile is affine: slot j's column is its base
    // plus the row's position in the tile
    const int32_t pos = (int32_t)(i % kTile);
#pragma unroll
    for (int j = 0; j < kSlots; ++j)
      f(j, (int64_t)j * n + i, base[(int64_t)j * C.ntiles] + pos);
  } else if (info == kTile8 && C.use8) {
#pragma unroll
    for (int g = 0; g < kSlotWords; ++g) {
      const uint32_t w = mload<STREAM>(C.off8 + (int64_t)g * n + i);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = 4 * g + b;
        if (j < kSlots)
          f(j, (int64_t)j * n + i,
            base[(int64_t)j * C.ntiles] + (int32_t)((w >> (8 * b)) & 0xffu));
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t k = (int64_t)j * n + i;
      f(j, k, base[(int64_t)j * C.ntiles] + (int32_t)mload<STREAM>(C.off + k));
    }
  }
}

// On a tile whose offset words are all affine (slot j's column is its base
// plus the row's position in the tile), slots 3m and 3m + 2 whose bases
// differ by one read x windows one element apart: on the 27-point grid the
// x - 1 and x + 1 neighbours, consecutive rows of one colour block.  Where
// that holds for all nine m, slot 3m + 2's x values are slot 3m's moved down
// one lane (ds_bpermute), and lane 63's is read on its own (one address for
// the whole wavefront), in place of nine 64-lane gathers.  Needs all 64
// lanes of the tile active (a full tile).  f(j, k, c, x_c) for every slot j
// in order.
template <class F>
__device__ inline void affine_pairs_row(const Col16 &C, int t, int64_t n, int64_t i,
                                        const double *x, F &&f) {
  const int32_t *__restrict__ base = C.base + t;
  const int32_t pos = (int32_t)(i % kTile);
#pragma unroll
  for (int m = 0; m < kSlots / 3; ++m) {
    const int32_t b0 = base[(int64_t)(3 * m) * C.ntiles];
    const int32_t b1 = base[(int64_t)(3 * m + 1) * C.ntiles];
    const double x0 = x[b0 + pos];
    const double x1 = x[b1 + pos];
    const double up = __shfl_down(x0, 1, 64);
    const double last = x[b0 + kTile];
    const double x2 = pos == kTile - 1 ? last : up;
    f(3 * m, (int64_t)(3 * m) * n + i, b0 + pos, x0);
    f(3 * m + 1, (int64_t)(3 * m + 1) * n + i, b1 + pos, x1);
    f(3 * m + 2, (int64_t)(3 * m + 2) * n + i, b0 + 1 + pos, x2);
  }
}

// Whether the wavefront of row i can take affine_pairs_row: an affine
// 8-bit tile all of whose 64 rows lie below `end`, with slot 3m + 2's base
// one above slot 3m's for every m (uniform per wavefront).
__device__ inline bool pairs_tile(const Col16 &C, int t, int info, int64_t end) {
  if (!(info == kTile8 && C.use8 && C.amask && C.amask[t] == (1u << kSlotWords) - 1u &&
        (int64_t)(t + 1) * kTile <= end))
    return false;
  const int32_t *__restrict__ base = C.base + t;
  bool ok = true;
#pragma unroll
  for (int m = 0; m < kSlots / 3; ++m)
    ok = ok && base[(int64_t)(3 * m + 2) * C.ntiles] == base[(int64_t)(3 * m) * C.ntiles] + 1;
  return ok;
}

template <bool STREAM, bool PAIRS>
__device__ inline double row_dot16(const Col16 &C, const int32_t *__restrict__ col,
                                   const double *__restrict__ val,
                                   const double *__restrict__ x, int64_t n,
                                   int64_t i, int64_t end) {
  const int t = tile_of(i);
  const int info = C.info[t];
  if (info == kTile32) return row_dot<STREAM>(col, val, x, n, i);
  double sum = 0.0;
  if (PAIRS && pairs_tile(C, t, info, end)) {
    affine_pairs_row(C, t, n, i, x, [&](int, int64_t k, int32_t, double xc) {
      sum += mload<STREAM>(val + k) * xc;
    });
    return sum;
  }
  narrow_row<STREAM>(C, t, info, n, i, [&](int, int64_t k, int32_t c) {
    sum += mload<STREAM>(val + k) * x[c];
  });
  return sum;
}

template <bool STREAM, bool PAIRS = false>
__device__ inline double gs_row16(const Col16 &C, int64_t i, int64_t n,
                                  const int32_t *__restrict__ col,
                                  const double *__restrict__ val,
                                  const double *__restrict__ r, double *x, int64_t end = 0) {
  const int t = tile_of(i);
  const int info = C.info[t];
  if (info == kTile32) return gs_row<STREAM>(i, n, col, val, r, x);
  double sum = r[i];
  double d = 0.0;
  if (PAIRS && pairs_tile(C, t, info, end)) {
    affine_pairs_row(C, t, n, i, x, [&](int, int64_t k, int32_t c, double xc) {
      const double v = mload<STREAM>(val + k);
      sum -= v * xc;
      if (c == i && v != 0.0) d = v;
    });
  } else {
    narrow_row<STREAM>(C, t, info, n, i, [&](int, int64_t k, int32_t c) {
      const double v = mload<STREAM>(val + k);
      sum -= v * x[c];
      if (c == i && v != 0.0) d = v;
    });
  }
  sum += x[i] * d;
  x[i] = sum / d;
  return sum / d;
}

// All rows of one colour (a contiguous device-order range) at once: rows of
// one colour share no matrix entry, so their updates are independent.
// FINEST only separates the finest level's launches (a distinct symbol for
// profiling); the code is the same.
template <bool STREAM, bool FINEST>
__global__ __launch_bounds__(kBlock) void k_gs_range(
    int64_t begin, int64_t end, int64_t n, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ r, double *x,
    bool xcd) {
  const int64_t b = xcd_block(blockIdx.x, gridDim.x, xcd);
  const int64_t i = begin + b * kBlock + threadIdx.x;
  if (i < end) gs_row<STREAM>(i, n, col, val, r, x);
}

// k_gs_range with 16-bit column ids (begin is a multiple of kTile).  BS
// threads per workgroup; W: the waves-per-EU floor given to the compiler
// (7 caps it at 72 VGPRs, where its own choice, 77, allows 6 waves per
// SIMD; 1 leaves it free).  The finest level's workgroup size and cap are
// tuning knobs ("gs_block", "gs_waves"); other levels use kBlock and 7.
template <bool STREAM, bool FINEST, int BS, int W, bool PAIRS = false>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(W))) void k_gs_range16(
    int64_t begin, int64_t end, int64_t n, Col16 C, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ r, double *x, bool xcd) {
  const int64_t i = begin + xcd_block(blockIdx.x, gridDim.x, xcd) * BS + threadIdx.x;
  if (i < end) gs_row16<STREAM, PAIRS>(C, i, n, col, val, r, x, end);
}

// k_gs_range16 that also stores the sum of r_i * (new x_i) over the rows of
// each wavefront with rows: wavefront k of the launch (rows begin + 64 k ..)
// writes dpart[k], whatever the workgroup size.  In the last sweep of a
// V-cycle every row is updated once, so these partials hold r.z (CG.cpp:110)
// without reading r and z again.  One partial per wavefront: a 64-lane
// shuffle tree, with no LDS and no workgroup barrier (a per-workgroup sum
// cost 40 us per launch at 512x512x640).  Under W = 6 the compiler fits 71
// VGPRs, which still run 7 waves per SIMD as k_gs_range16 does, and keeps
// k_gs_range16's memory waits; under W = 7 (70 VGPRs) it waited more often
// and the step took 12% longer.  Uncapped it takes 120 (4 waves).
template <bool STREAM, int BS, int W, bool PAIRS = false>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(W))) void k_gs_rz16(
    int64_t begin, int64_t end, int64_t n, Col16 C, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ r, double *x, bool xcd,
    double *__restrict__ dpart) {
  const int64_t b = xcd_block(blockIdx.x, gridDim.x, xcd);
  const int64_t i = begin + b * BS + threadIdx.x;
  double s = 0.0;
  if (i < end) s = r[i] * gs_row16<STREAM, PAIRS>(C, i, n, col, val, r, x, end);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  // the wavefront's index in the launch, formed after the row so that it
  // holds no register during it
  const int64_t w = b * (BS / 64) + (threadIdx.x >> 6);
  if ((threadIdx.x & 63) == 0 && begin + w * 64 < end) dpart[w] = s;
}

// The first forward sweep of a V-cycle level from x = 0 (ComputeMG_ref.cpp:
// 42-47 zeroes x before the pre-smoother).  In the colour step over device
// rows [begin, end), the only x values the reference's row update reads
// that are not 0 are those of rows updated earlier in this sweep: device
// rows below `begin`.  So a slot's values are read only when Level::zmask
// marks it for the tile (some row has a nonzero there with a column below
// its colour's first row; uniform per tile), and x only below `begin`.
// Skipped terms are v * 0 and the diagonal term is x_i * d = 0 * d:
// leaving them out changes no nonzero sum (only, for a zero sum, the sign
// of the zero), so x agrees with the reference's sweep of a zero-filled x,
// and x needs no fill: every x this sweep reads was written by it.  The
// diagonal comes from the diagonal array (the stored entry whose column is
// the row, as matrixDiagonal).
//
// Every slot still issues its loads, as the full colour step does, so the
// loads of a row stay in flight together; the skipping is done by the
// buffer unit's range check: a load through a buffer view at a byte offset
// at or beyond the view's size returns 0 and reads no memory.  A dead
// slot's views have size 0; the x view of a live slot ends at `begin`.
typedef __amdgpu_buffer_rsrc_t BufView;
constexpr int kBufWord3 = 0x00020000;  // buffer descriptor word 3 for gfx9 raw loads

__device__ inline BufView buf_view(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kBufWord3);
}
template <bool STREAM>
__device__ inline double buf_f64(BufView v, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(v, off, 0, STREAM ? 2 : 0));
}
template <bool STREAM>
__device__ inline uint32_t buf_u32(BufView v, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(v, off, 0, STREAM ? 2 : 0);
}
template <bool STREAM>
__device__ inline uint32_t buf_u16(BufView v, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b16(v, off, 0, STREAM ? 2 : 0);
}

// Row i (lane `lane` of tile t, whose first row is tb) of a zero-x colour
// step; zm is the tile's Level::zmask.  n * 8 < 2^32 (x offsets).  The
// tile's class (32-bit ids, all-affine 8-bit offsets, 8-bit, 16-bit) is
// decided once, outside the slot loops, so each class's loop is straight
// code whose loads the compiler can issue together.
template <bool STREAM>
__device__ inline void gs_row_zero(const Col16 &C, uint32_t zm, int64_t begin, int64_t i,
                                   int64_t n, const int32_t *__restrict__ col,
                                   const double *__restrict__ val,
                                   const double *__restrict__ diag,
                                   const double *__restrict__ r, double *x) {
  const int t = tile_of(i);
  const int info = C.info[t];
  const int64_t tb = (int64_t)t * kTile;
  const uint32_t lane = (uint32_t)(i - tb);
  const uint32_t xlim = (uint32_t)begin * 8u;
  const int32_t *__restrict__ base = C.base + t;
  double sum = r[i];
  // slot j's term, its value and x read through views that are empty for a
  // dead slot and end at `begin` for x
  auto term = [&](int j, int32_t c) {
    const bool live = (zm >> j) & 1u;
    const double v =
        buf_f64<STREAM>(buf_view(val + (int64_t)j * n + tb, live ? 512u : 0u), lane * 8u);
    const double xc = buf_f64<false>(buf_view(x, live ? xlim : 0u), (uint32_t)c * 8u);
    sum -= v * xc;
  };
  if (info == kTile32) {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const bool live = (zm >> j) & 1u;
      term(j, (int32_t)buf_u32<STREAM>(buf_view(col + (int64_t)j * n + tb, live ? 256u : 0u),
                                       lane * 4u));
    }
  } else if (info == kTile8 && C.use8 && C.amask &&
             C.amask[t] == (1u << kSlotWords) - 1u) {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) term(j, base[(int64_t)j * C.ntiles] + (int32_t)lane);
  } else if (info == kTile8 && C.use8) {
#pragma unroll
    for (int g = 0; g < kSlotWords; ++g) {
      const bool any = (zm >> (4 * g)) & 0xfu;
      const uint32_t w = buf_u32<STREAM>(
          buf_view(C.off8 + (int64_t)g * n + tb, any ? 256u : 0u), lane * 4u);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = 4 * g + b;
        if (j < kSlots)
          term(j, base[(int64_t)j * C.ntiles] + (int32_t)((w >> (8 * b)) & 0xffu));
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const bool live = (zm >> j) & 1u;
      term(j, base[(int64_t)j * C.ntiles] +
                  (int32_t)buf_u16<STREAM>(buf_view(C.off + (int64_t)j * n + tb,
                                                    live ? 128u : 0u),
                                           lane * 2u));
    }
  }
  const double d = mload<STREAM>(diag + i);
  sum += 0.0 * d;
  x[i] = sum / d;
}

constexpr int kLiveChunk = 8;  // live slots per chunk (8 measured faster than 16)

// gs_row_zero over the live slots only ("zero_live"): Level::zmask's set
// bits in ascending slot order, CH per chunk, with the chunk's unused
// entries read through empty views (a * 0).  The dead slots' loads are not
// issued at all: gs_row_zero issues every slot's loads, and a step with few
// live slots then cost as many instructions as a full one.  CLS: the tile
// class (0 32-bit ids, 1 all-affine 8-bit, 2 8-bit, 3 16-bit), decided once
// per tile outside the loop.
template <bool STREAM, int CLS, int CH>
__device__ inline double zero_live_sum(const Col16 &C, uint32_t zm, int t, int64_t tb,
                                       uint32_t lane, uint32_t xlim, int64_t n,
                                       const int32_t *__restrict__ col,
                                       const double *__restrict__ val, const double *x,
                                       double sum) {
  const int32_t *__restrict__ base = C.base + t;
  uint32_t mm = zm;
  while (mm) {
    int js[CH];
    bool ok[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      ok[q] = mm != 0u;
      js[q] = ok[q] ? __builtin_ctz(mm) : 0;
      mm &= mm - 1u;
    }
    // the chunk's index words and values, then its x values, then the sums
    uint32_t w[CH];
    double v[CH], xc[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int j = js[q];
      if (CLS == 0)
        w[q] = buf_u32<STREAM>(buf_view(col + (int64_t)j * n + tb, ok[q] ? 256u : 0u), lane * 4u);
      else if (CLS == 2)
        w[q] = buf_u32<STREAM>(buf_view(C.off8 + (int64_t)(j >> 2) * n + tb, ok[q] ? 256u : 0u),
                               lane * 4u);
      else if (CLS == 3)
        w[q] = buf_u16<STREAM>(buf_view(C.off + (int64_t)j * n + tb, ok[q] ? 128u : 0u),
                               lane * 2u);
      v[q] = buf_f64<STREAM>(buf_view(val + (int64_t)j * n + tb, ok[q] ? 512u : 0u), lane * 8u);
    }
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int j = js[q];
      int32_t c;
      if (CLS == 0)
        c = (int32_t)w[q];
      else if (CLS == 1)
        c = base[(int64_t)j * C.ntiles] + (int32_t)lane;
      else if (CLS == 2)
        c = base[(int64_t)j * C.ntiles] + (int32_t)((w[q] >> (8 * (j & 3))) & 0xffu);
      else
        c = base[(int64_t)j * C.ntiles] + (int32_t)w[q];
      xc[q] = buf_f64<false>(buf_view(x, ok[q] ? xlim : 0u), (uint32_t)c * 8u);
    }
#pragma unroll
    for (int q = 0; q < CH; ++q) sum -= v[q] * xc[q];
  }
  return sum;
}

template <bool STREAM, int CH>
__device__ inline void gs_row_zero_live(const Col16 &C, uint32_t zm, int64_t begin, int64_t i,
                                        int64_t n, const int32_t *__restrict__ col,
                                        const double *__restrict__ val,
                                        const double *__restrict__ diag,
                                        const double *__restrict__ r, double *x) {
  // every wavefront's rows lie in one tile (colour blocks start on tiles in
  // this layout, and workgroups are whole wavefronts), so the tile and its
  // mask are uniform: the slot loop runs on scalar registers
  const int t = __builtin_amdgcn_readfirstlane(tile_of(i));
  zm = __builtin_amdgcn_readfirstlane(zm);
  const int info = C.info[t];
  const int64_t tb = (int64_t)t * kTile;
  const uint32_t lane = (uint32_t)(i - tb);
  const uint32_t xlim = (uint32_t)begin * 8u;
  double sum = r[i];
  if (info == kTile32)
    sum = zero_live_sum<STREAM, 0, CH>(C, zm, t, tb, lane, xlim, n, col, val, x, sum);
  else if (info == kTile8 && C.use8 && C.amask && C.amask[t] == (1u << kSlotWords) - 1u)
    sum = zero_live_sum<STREAM, 1, CH>(C, zm, t, tb, lane, xlim, n, col, val, x, sum);
  else if (info == kTile8 && C.use8)
    sum = zero_live_sum<STREAM, 2, CH>(C, zm, t, tb, lane, xlim, n, col, val, x, sum);
  else
    sum = zero_live_sum<STREAM, 3, CH>(C, zm, t, tb, lane, xlim, n, col, val, x, sum);
  const double d = mload<STREAM>(diag + i);
  sum += 0.0 * d;
  x[i] = sum / d;
}

// LIVE: the live slots only (gs_row_zero_live), else every slot (gs_row_zero)
template <bool STREAM, bool FINEST, int BS, bool LIVE = false>
__global__ __launch_bounds__(BS) void k_gs_zero16(
    int64_t begin, int64_t end, int64_t n, Col16 C, const uint32_t *__restrict__ zmask,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ diag, const double *__restrict__ r, double *x, bool xcd) {
  const int64_t i = begin + xcd_block(blockIdx.x, gridDim.x, xcd) * BS + threadIdx.x;
  if (i >= end) return;
  if (LIVE)
    gs_row_zero_live<STREAM, kLiveChunk>(C, zmask[tile_of(i)], begin, i, n, col, val, diag, r, x);
  else
    gs_row_zero<STREAM>(C, zmask[tile_of(i)], begin, i, n, col, val, diag, r, x);
}

// Two adjacent rows per thread: the matrix is read with 16-byte (values) and
// 8-byte (column ids) loads per lane.  Needs an even level size and an even
// range start, so that slot j of the pair is 16-byte aligned.
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef int i32x2 __attribute__((ext_vector_type(2)));

template <bool STREAM>
__device__ inline void gs_row_pair(int64_t i, int64_t n,
                                   const int32_t *__restrict__ col,
                                   const double *__restrict__ val,
                                   const double *__restrict__ r, double *x) {
  const f64x2 rr = *reinterpret_cast<const f64x2 *>(r + i);
  double s0 = rr.x, s1 = rr.y, d0 = 0.0, d1 = 0.0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t k = (int64_t)j * n + i;
    const f64x2 v = mload<STREAM>(reinterpret_cast<const f64x2 *>(val + k));
    const i32x2 c = mload<STREAM>(reinterpret_cast<const i32x2 *>(col + k));
    s0 -= v.x * x[c.x];
    s1 -= v.y * x[c.y];
    if (c.x == i && v.x != 0.0) d0 = v.x;
    if (c.y == i + 1 && v.y != 0.0) d1 = v.y;
  }
  s0 += x[i] * d0;
  s1 += x[i + 1] * d1;
  x[i] = s0 / d0;
  x[i + 1] = s1 / d1;
}

template <bool STREAM, bool FINEST>
__global__ __launch_bounds__(kBlock) void k_gs_range2(
    int64_t begin, int64_t end, int64_t n, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ r, double *x) {
  const int64_t i = begin + 2 * gtid();
  if (i + 1 < end)
    gs_row_pair<STREAM>(i, n, col, val, r, x);
  else if (i < end)
    gs_row<STREAM>(i, n, col, val, r, x);
}

// Multicolour sweep in one launch, as a wavefront over natural-row slabs
// (see build_wave for the task order).  The grid is co-resident (cooperative
// launch) and workgroup g runs tasks g, g+G, g+2G, ... in list order, so the
// earliest unfinished task always has its inputs done: the waits end.  A wait
// that exceeds kSpinLimit polls (seconds) sets *err and gives up, as do all
// later ones; the sweep then finishes with wrong values instead of hanging.
//
// Visibility of x across workgroups (and XCDs) without cache maintenance:
// inside this kernel x is only accessed with device-scope (agent) atomic
// loads and stores, which are served at the device-coherent level (the
// Infinity Cache / HBM) rather than by the per-XCD L2.  Every wave waits for
// its stores to complete (s_waitcnt 0) before the barrier after which thread
// 0 bumps the tile's counter, and readers issue their x loads only after
// thread 0 has seen the counter.  Agent-scope fences are not used: each one
// writes back / invalidates the XCD's whole L2.
constexpr unsigned kSpinLimit = 1u << 22;

__device__ inline double xload(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void xstore(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// gs_row inside the wavefront: x of rows already updated in this launch
// (colours of earlier phases: device rows below `lo` in the forward sweep,
// at or above `hi` in the backward sweep, [lo, hi) being this row's colour
// block) is read device-coherently; every other x value cannot have changed
// since the launch began and is read through the caches.  The new x is
// stored device-coherently.
template <bool STREAM, int DIR>
__device__ inline void gs_row_wave(int64_t i, int64_t n, int64_t lo, int64_t hi,
                                   const int32_t *__restrict__ col,
                                   const double *__restrict__ val,
                                   const double *__restrict__ r, double *x) {
  double sum = r[i];
  double d = 0.0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t k = (int64_t)j * n + i;
    const double v = mload<STREAM>(val + k);
    const int32_t c = mload<STREAM>(col + k);
    const bool fresh = DIR == 0 ? c < lo : c >= hi;
    const double xc = fresh ? xload(x + c) : x[c];
    sum -= v * xc;
    if (c == i && v != 0.0) d = v;
  }
  sum += x[i] * d;
  xstore(x + i, sum / d);
}

template <bool STREAM, bool FINEST, int DIR>
__global__ __launch_bounds__(kBlock) void k_gs_wave(
    const int32_t *__restrict__ tasks, int64_t ntasks,
    const int64_t *__restrict__ tile_ptr, int S, int C, int chunk_rows,
    unsigned *done, int *err, int64_t n, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ r, double *x) {
  for (int64_t k = blockIdx.x; k < ntasks; k += gridDim.x) {
    const int tile = tasks[2 * k];
    const int chunk = tasks[2 * k + 1];
    const int p = tile / S, sl = tile % S;
    if (threadIdx.x == 0 && p > 0) {
      const int cq = DIR == 0 ? p - 1 : C - p;  // colour of phase p-1
      const int64_t *tq = tile_ptr + (int64_t)cq * (S + 1);
      const int lo = sl > 0 ? sl - 1 : 0, hi = sl + 1 < S ? sl + 1 : S - 1;
      for (int s2 = lo; s2 <= hi; ++s2) {
        const unsigned need =
            (unsigned)((tq[s2 + 1] - tq[s2] + chunk_rows - 1) / chunk_rows);
        unsigned spins = 0;
        while (__hip_atomic_load(&done[(p - 1) * S + s2], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) < need) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins > kSpinLimit ||
              __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            atomicOr(err, 1);
            break;
          }
        }
      }
    }
    __syncthreads();
    const int c = DIR == 0 ? p : C - 1 - p;
    const int64_t *tp = tile_ptr + (int64_t)c * (S + 1);
    const int64_t cb = tp[0], ce = tp[S];  // this colour's device rows
    const int64_t b0 = tp[sl] + (int64_t)chunk * chunk_rows;
    const int64_t e0 = min(b0 + chunk_rows, tp[sl + 1]);
    for (int64_t i = b0 + threadIdx.x; i < e0; i += kBlock)
      gs_row_wave<STREAM, DIR>(i, n, cb, ce, col, val, r, x);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's x stores are complete
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_fetch_add(&done[p * S + sl], 1u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ __launch_bounds__(kBlock) void k_zero_u32(unsigned *a, int64_t n) {
  const int64_t i = gtid();
  if (i < n) a[i] = 0u;
}

// All rows of one level of the lexicographic schedule (a list of rows).
__global__ __launch_bounds__(kBlock) void k_gs_list(
    const int32_t *__restrict__ rows, int64_t count, int64_t n,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ r, double *x) {
  const int64_t k = gtid();
  if (k < count) gs_row<false>(rows[k], n, col, val, r, x);
}

// ComputeRestriction_ref.cpp:51
__global__ __launch_bounds__(kBlock) void k_restrict(
    int64_t nc, const int32_t *__restrict__ f2c, const double *__restrict__ rf,
    const double *__restrict__ Axf, double *__restrict__ rc) {
  const int64_t c = gtid();
  if (c < nc) {
    const int32_t f = f2c[c];
    rc[c] = rf[f] - Axf[f];
  }
}

// ComputeMG_ref.cpp:49-51 fused: the fine-level product only at injected rows.
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_resid_restrict(
    int64_t nc, int64_t n, const int32_t *__restrict__ f2c,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ rf, const double *__restrict__ xf,
    double *__restrict__ rc, bool xcd) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x, xcd) * kBlock + threadIdx.x;
  if (c < nc) {
    const int64_t f = f2c[c];
    rc[c] = rf[f] - row_dot<STREAM>(col, val, xf, n, f);
  }
}

// The same in fine order (Level::ftile, rperm): thread t computes fine row
// fine_row(ftile, t), so a wavefront reads whole tiles of the matrix (narrow
// column ids when C16), and stores the coarse row rperm[t].
// PAIRS: row_dot16's x-pair path on wavefronts whose 64 transfer rows are
// all present (each is then a whole fine tile).
template <bool STREAM, bool C16, bool PAIRS = false>
__global__ __launch_bounds__(kBlock) void k_resid_restrict_f(
    int64_t nc, const int32_t *__restrict__ ftile, const int32_t *__restrict__ rperm, int64_t n,
    Col16 C, const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ rf, const double *__restrict__ xf, double *__restrict__ rc) {
  const int64_t t = gtid();
  if (t >= nc) return;
  const int64_t f = fine_row(ftile, t);
  const int64_t end = (t / kTile + 1) * kTile <= nc ? n : 0;
  const double ax = C16 ? row_dot16<STREAM, PAIRS>(C, col, val, xf, n, f, end)
                        : row_dot<STREAM>(col, val, xf, n, f);
  rc[rperm[t]] = rf[f] - ax;
}

// ComputeProlongation_ref.cpp:49 (f2c is injective: no write conflicts)
__global__ __launch_bounds__(kBlock) void k_prolong_f(int64_t nc,
                                                      const int32_t *__restrict__ ftile,
                                                      const int32_t *__restrict__ rperm,
                                                      const double *__restrict__ xc,
                                                      double *__restrict__ xf) {
  const int64_t t = gtid();
  if (t < nc) xf[fine_row(ftile, t)] += xc[rperm[t]];
}

__global__ __launch_bounds__(kBlock) void k_prolong(
    int64_t nc, const int32_t *__restrict__ f2c, const double *__restrict__ xc,
    double *__restrict__ xf) {
  const int64_t c = gtid();
  if (c < nc) xf[f2c[c]] += xc[c];
}

__global__ __launch_bounds__(kBlock) void k_replace_diag(
    int64_t n, const uint8_t *__restrict__ dslot, const double *__restrict__ d,
    double *__restrict__ diag, double *__restrict__ val) {
  const int64_t i = gtid();
  if (i < n) {
    const double v = d[i];
    diag[i] = v;
    val[(int64_t)dslot[i] * n + i] = v;
  }
}

// ---------------------------------------------------------------------------
// Deterministic grid reductions
// ---------------------------------------------------------------------------
// Each workgroup stores its partial; the last one to finish (ticket) sums the
// partials in index order and applies the finishing step.  The ticket is
// reset by that workgroup, so the next launch starts from zero.
__device__ inline void grid_finish(double s, double *partials, unsigned *ticket,
                                   double *scalars, double *out, int finish) {
  __shared__ int is_last;
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = s;
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  double v = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
    v += __hip_atomic_load(&partials[b], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  v = block_sum(v);
  if (threadIdx.x == 0) {
    *ticket = 0u;
    if (finish == kFinishStore) {
      *out = v;
    } else if (finish == kFinishBeta) {  // CG.cpp:112-114
      const double old = scalars[kRtz];
      scalars[kOldRtz] = old;
      scalars[kRtz] = v;
      scalars[kBeta] = v / old;
    } else {  // kFinishAlpha, CG.cpp:119-120
      scalars[kPAp] = v;
      scalars[kAlpha] = scalars[kRtz] / v;
    }
  }
}

// p.Ap from the per-chunk partials of k_spmv_dot_c / k_spmv_dot_hc (or r.z
// from those of k_gs_rz16), in a launch of its own (the launch boundary orders the partials' stores before
// these loads): thread t of workgroup w sums chunks w*per + t, + kBlock, ...
// in order, then block_sum and grid_finish, so the order is fixed.
__global__ __launch_bounds__(kBlock) void k_chunk_dot(int64_t nchunks,
                                                      const double *__restrict__ cpart,
                                                      double *partials, unsigned *ticket,
                                                      double *scalars, int finish,
                                                      double *out = nullptr) {
  const int64_t per = (nchunks + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < nchunks ? lo + per : nchunks;
  double s = 0.0;
  for (int64_t q = lo + threadIdx.x; q < hi; q += kBlock) s += cpart[q];
  s = block_sum(s);
  grid_finish(s, partials, ticket, scalars, out, finish);
}

constexpr int64_t kStreamRows = (int64_t)1 << 24;
inline bool stream_vec(int64_t n) { return n >= kStreamRows; }

inline unsigned reduce_grid(int64_t n) {
  const int64_t g = (n + kBlock - 1) / kBlock;
  return (unsigned)(g < kReduceBlocks ? (g < 1 ? 1 : g) : kReduceBlocks);
}

// Workgroups of k_chunk_dot over n partials: about 8 per thread, at most
// 256 workgroups.  Each workgroup of the last-workgroup finish passes a
// device-scope fence; with reduce_grid's up to 4096 workgroups the fences,
// not the few megabytes of partials, set the time (55-93 us per launch at
// 512x512x640).
inline unsigned chunk_grid(int64_t n) {
  const int64_t g = (n + 8 * kBlock - 1) / (8 * kBlock);
  return (unsigned)(g < 256 ? (g < 1 ? 1 : g) : 256);
}

// The grid-stride loops below visit i = gtid(), gtid() + stride, ... in that
// order and add into one running sum in that order; the main loop loads
// four of those rows before it uses any, so four rows' loads are in flight
// per wavefront (the sums are the same as one row at a time).
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const double *x,
                                                 const double *y,
                                                 double *partials,
                                                 unsigned *ticket,
                                                 double *scalars, double *out,
                                                 int finish) {
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = gtid();
  for (; i + 3 * stride < n; i += 4 * stride) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = mload<STREAM>(x + i + u * stride);
      b[u] = mload<STREAM>(y + i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) s += a[u] * b[u];
  }
  for (; i < n; i += stride) s += mload<STREAM>(x + i) * mload<STREAM>(y + i);
  s = block_sum(s);
  grid_finish(s, partials, ticket, scalars, out, finish);
}

// The dot's first pass (ComputeDotProduct_ref.cpp:46-80 restated with a
// fixed tree): workgroup b takes elements [b * kDotPer, ...), its thread q
// the elements b * kDotPer + q + e * kBlock (e < kDotE, all loads issued
// before the sum), in that order, then block_sum; the workgroup's partial is
// stored plainly and k_chunk_dot sums the partials in a second launch.  The
// grid-stride k_dot above finishes in its last workgroup, behind a
// device-scope fence in each of its 4096 workgroups: 0.47 ms of loads plus
// the fences took 0.62 ms on two vectors of 512x512x640 doubles, this
// pass and k_chunk_dot 0.39 ms.
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_dot_rows(int64_t n, const double *x, const double *y,
                                                     double *__restrict__ parts) {
  const int64_t t0 = (int64_t)blockIdx.x * kDotPer + threadIdx.x;
  double a[kDotE], b[kDotE];
#pragma unroll
  for (int e = 0; e < kDotE; ++e) {
    const int64_t t = t0 + (int64_t)e * kBlock;
    a[e] = t < n ? mload<STREAM>(x + t) : 0.0;
    b[e] = t < n ? mload<STREAM>(y + t) : 0.0;
  }
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < kDotE; ++e) s += a[e] * b[e];
  s = block_sum(s);
  if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// Ap = A p and p.Ap (CG.cpp:118-120) in one pass.  A persistent grid (a
// multiple of 8 workgroups): XCD x (= blockIdx % 8) owns the x-th eighth of
// the colour-interleaved chunk list and its workgroups stride through it, so
// both the partial sums and each XCD's gather window are fixed.
template <bool STREAM, bool C16>
__global__ __launch_bounds__(kBlock) void k_spmv_dot(
    int64_t nchunks, const int64_t *__restrict__ chunks, int64_t n, Col16 C,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ p, double *__restrict__ Ap, double *partials,
    unsigned *ticket, double *scalars) {
  const int64_t x = blockIdx.x % 8, slot = blockIdx.x / 8, per = gridDim.x / 8;
  const int64_t lo = x * nchunks / 8, hi = (x + 1) * nchunks / 8;
  double s = 0.0;
  for (int64_t w = lo + slot; w < hi; w += per) {
    const int64_t i = chunks[2 * w] + threadIdx.x;
    if (i < chunks[2 * w + 1]) {
      const double y = C16 ? row_dot16<STREAM>(C, col, val, p, n, i)
                           : row_dot<STREAM>(col, val, p, n, i);
      Ap[i] = y;
      s += p[i] * y;
    }
  }
  s = block_sum(s);
  grid_finish(s, partials, ticket, scalars, nullptr, kFinishAlpha);
}

// Ap = A p and p.Ap for the mixed-precision outer iteration, reading the
// FP32 V-cycle's packed words (16-bit column offset above the FP16 value)
// for the off-diagonal values and the FP64 diagonal array for the diagonal,
// in slot order with FP64 products and sums: the FP64 SpMV whenever the
// off-diagonal values are exact in FP16 (they are -1 here), at 4 instead of
// 9 matrix bytes per entry.  32-bit tiles read the FP64 ELL.
// k_spmv_dot with one workgroup per row chunk (logical chunk b of the
// launch's workgroup, XCD-remapped when xcd): chunk b's p.Ap partial goes to
// cpart[b] for k_chunk_dot.
template <bool STREAM, bool C16, bool PAIRS = false>
__global__ __launch_bounds__(kBlock) void k_spmv_dot_c(
    const int64_t *__restrict__ chunks, int64_t n, Col16 C, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ p, double *__restrict__ Ap,
    double *__restrict__ cpart, bool xcd) {
  const int64_t b = xcd_block(blockIdx.x, gridDim.x, xcd);
  const int64_t i = chunks[2 * b] + threadIdx.x;
  const int64_t end = chunks[2 * b + 1];
  double s = 0.0;
  if (i < end) {
    const double y = C16 ? row_dot16<STREAM, PAIRS>(C, col, val, p, n, i, end)
                         : row_dot<STREAM>(col, val, p, n, i);
    Ap[i] = y;
    s = p[i] * y;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) cpart[b] = s;
}

// Row i of A p in FP64 from the FP16 off-diagonal values (packed words)
// and the FP64 diagonal, in slot order; 32-bit tiles read the FP64 matrix.
template <bool STREAM>
__device__ inline double row_dot_h(const Col16 &C, const uint32_t *__restrict__ vc,
                                   const uint8_t *__restrict__ dslot,
                                   const double *__restrict__ diag,
                                   const int32_t *__restrict__ col,
                                   const double *__restrict__ val,
                                   const double *__restrict__ p, int64_t n, int64_t i) {
  const int t = tile_of(i);
  if (C.info[t] == kTile32) return row_dot<STREAM>(col, val, p, n, i);
  const int32_t *__restrict__ base = C.base + t;
  const int ds = mload<STREAM>(dslot + i);
  const double di = mload<STREAM>(diag + i);
  double y = 0.0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const uint32_t wd = mload<STREAM>(vc + (int64_t)j * n + i);
    const int32_t c = base[(int64_t)j * C.ntiles] + (int32_t)(wd >> 16);
    const double a = j == ds ? di : (double)__builtin_bit_cast(_Float16, (uint16_t)(wd & 0xffffu));
    y += a * p[c];
  }
  return y;
}

// row_dot_h on an 8-bit tile from the FP16 value pairs (vp, kPairWords words)
// and the 8-bit offsets (none loaded where the tile's amask bits say affine):
// the same FP16 values, columns and slot order, 2 to 3 bytes per slot in
// place of 4.  Other tiles take row_dot_h.
template <bool STREAM>
__device__ inline double row_dot_hp(const Col16 &C, const uint32_t *__restrict__ vp,
                                    const uint32_t *__restrict__ vc,
                                    const uint8_t *__restrict__ dslot,
                                    const double *__restrict__ diag,
                                    const int32_t *__restrict__ col,
                                    const double *__restrict__ val,
                                    const double *__restrict__ p, int64_t n, int64_t i) {
  const int t = tile_of(i);
  if (C.info[t] != kTile8 || !C.use8)
    return row_dot_h<STREAM>(C, vc, dslot, diag, col, val, p, n, i);
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
  const unsigned am = C.amask ? C.amask[t] : 0u;
  uint32_t o8[kSlotWords], pv[kPairWords];
#pragma unroll
  for (int g = 0; g < kSlotWords; ++g) o8[g] = off8_word<STREAM>(C.off8, am, g, n, i);
#pragma unroll
  for (int w = 0; w < kPairWords; ++w) pv[w] = mload<STREAM>(vp + (int64_t)w * n + i);
  const int ds = mload<STREAM>(dslot + i);
  const double di = mload<STREAM>(diag + i);
  double y = 0.0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int32_t c = base[j] + (int32_t)((o8[j / 4] >> (8 * (j % 4))) & 0xffu);
    const uint16_t h = (uint16_t)(pv[j / 2] >> (16 * (j % 2)));
    const double a = j == ds ? di : (double)__builtin_bit_cast(_Float16, h);
    y += a * p[c];
  }
  return y;
}

// k_spmv_dot_h with one workgroup per row chunk; partials as k_spmv_dot_c.
// vp non-null: 8-bit tiles take row_dot_hp.
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_spmv_dot_hc(
    const int64_t *__restrict__ chunks, int64_t n, Col16 C, const uint32_t *__restrict__ vc,
    const uint32_t *__restrict__ vp, const uint8_t *__restrict__ dslot,
    const double *__restrict__ diag, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ p, double *__restrict__ Ap,
    double *__restrict__ cpart, bool xcd) {
  const int64_t b = xcd_block(blockIdx.x, gridDim.x, xcd);
  const int64_t i = chunks[2 * b] + threadIdx.x;
  double s = 0.0;
  if (i < chunks[2 * b + 1]) {
    const double y = vp ? row_dot_hp<STREAM>(C, vp, vc, dslot, diag, col, val, p, n, i)
                        : row_dot_h<STREAM>(C, vc, dslot, diag, col, val, p, n, i);
    mstore<STREAM>(Ap + i, y);
    s = p[i] * y;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) cpart[b] = s;
}

template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_spmv_dot_h(
    int64_t nchunks, const int64_t *__restrict__ chunks, int64_t n, Col16 C,
    const uint32_t *__restrict__ vc, const uint8_t *__restrict__ dslot,
    const double *__restrict__ diag, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ p, double *__restrict__ Ap,
    double *partials, unsigned *ticket, double *scalars) {
  const int64_t xq = blockIdx.x % 8, slot = blockIdx.x / 8, per = gridDim.x / 8;
  const int64_t lo = xq * nchunks / 8, hi = (xq + 1) * nchunks / 8;
  double s = 0.0;
  for (int64_t w = lo + slot; w < hi; w += per) {
    const int64_t i = chunks[2 * w] + threadIdx.x;
    if (i < chunks[2 * w + 1]) {
      const double y = row_dot_h<STREAM>(C, vc, dslot, diag, col, val, p, n, i);
      mstore<STREAM>(Ap + i, y);
      s += p[i] * y;
    }
  }
  s = block_sum(s);
  grid_finish(s, partials, ticket, scalars, nullptr, kFinishAlpha);
}

// x += alpha p; r -= alpha Ap; r.r (CG.cpp:121-124).  R32: also writes r
// rounded to FP32, the input of the mixed-precision V-cycle.  WITH_X false:
// r and r.r only (x's step is taken by the next k_update_xp).
template <bool R32, bool STREAM, bool WITH_X>
__global__ __launch_bounds__(kBlock) void k_update_xr(
    int64_t n, double *__restrict__ x, double *__restrict__ r,
    const double *__restrict__ p, const double *__restrict__ Ap,
    float *__restrict__ r32, double *partials, unsigned *ticket, double *scalars) {
  const double alpha = scalars[kAlpha];
  const double malpha = -alpha;
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = gtid();
  for (; i + 3 * stride < n; i += 4 * stride) {
    double xv[4], pv[4], rv[4], av[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = i + u * stride;
      if (WITH_X) {
        xv[u] = mload<STREAM>(x + k);
        pv[u] = mload<STREAM>(p + k);
      }
      rv[u] = mload<STREAM>(r + k);
      av[u] = mload<STREAM>(Ap + k);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = i + u * stride;
      if (WITH_X) mstore<STREAM>(x + k, xv[u] + alpha * pv[u]);
      const double ri = rv[u] + malpha * av[u];
      mstore<STREAM>(r + k, ri);
      if (R32) mstore<STREAM>(r32 + k, (float)ri);
      s += ri * ri;
    }
  }
  for (; i < n; i += stride) {
    if (WITH_X) mstore<STREAM>(x + i, mload<STREAM>(x + i) + alpha * mload<STREAM>(p + i));
    const double ri = mload<STREAM>(r + i) + malpha * mload<STREAM>(Ap + i);
    mstore<STREAM>(r + i, ri);
    if (R32) mstore<STREAM>(r32 + i, (float)ri);
    s += ri * ri;
  }
  s = block_sum(s);
  grid_finish(s, partials, ticket, scalars, &scalars[kNormr2], kFinishStore);
}

// k_update_xr with one row per thread: each workgroup stores its r.r
// partial with a plain store (no ticket), and k_chunk_dot sums the partials
// in a fixed order in a second launch ("xr_rows").  At 512x512x640 the
// update without x takes 671 us this way against 909 us in the grid-stride
// form, which finishes r.r in its last workgroup (same job).
template <bool R32, bool STREAM, bool WITH_X>
__global__ __launch_bounds__(kBlock) void k_update_xr_rows(
    int64_t n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
    const double *__restrict__ Ap, float *__restrict__ r32, const double *__restrict__ scalars,
    double *__restrict__ part) {
  const int64_t i = gtid();
  double s = 0.0;
  if (i < n) {
    const double alpha = scalars[kAlpha];
    const double malpha = -alpha;
    if (WITH_X) mstore<STREAM>(x + i, mload<STREAM>(x + i) + alpha * mload<STREAM>(p + i));
    const double ri = mload<STREAM>(r + i) + malpha * mload<STREAM>(Ap + i);
    mstore<STREAM>(r + i, ri);
    if (R32) mstore<STREAM>(r32 + i, (float)ri);
    s = ri * ri;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// CG.cpp:121's step of the previous iteration, x += alpha p, taken here
// with CG.cpp:115's p = z + beta p: one pass reads x, p and z and writes x
// and p, where separate updates read p twice.  alpha is still the previous
// iteration's (this iteration's comes with its p.Ap, after this kernel).
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_update_xp(int64_t n, double *__restrict__ x,
                                                       const double *__restrict__ z,
                                                       double *__restrict__ p,
                                                       const double *__restrict__ scalars) {
  const int64_t i = gtid();
  if (i < n) {
    const double pi = mload<STREAM>(p + i);
    mstore<STREAM>(x + i, mload<STREAM>(x + i) + scalars[kAlpha] * pi);
    mstore<STREAM>(p + i, mload<STREAM>(z + i) + scalars[kBeta] * pi);
  }
}

// k_update_xp with z in FP32, widened exactly: the mixed CG's p update
// (CG.cpp:115) with the previous iteration's deferred x step (:121).
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_update_xp32(int64_t n, double *__restrict__ x,
                                                         const float *__restrict__ z,
                                                         double *__restrict__ p,
                                                         const double *__restrict__ scalars) {
  const int64_t i = gtid();
  if (i < n) {
    const double pi = mload<STREAM>(p + i);
    mstore<STREAM>(x + i, mload<STREAM>(x + i) + scalars[kAlpha] * pi);
    mstore<STREAM>(p + i, (double)mload<STREAM>(z + i) + scalars[kBeta] * pi);
  }
}

// x += alpha p alone (the last iteration's deferred step).
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_step_x(int64_t n, double *__restrict__ x,
                                                    const double *__restrict__ p,
                                                    const double *__restrict__ scalars) {
  const int64_t i = gtid();
  if (i < n) mstore<STREAM>(x + i, mload<STREAM>(x + i) + scalars[kAlpha] * mload<STREAM>(p + i));
}

// r.z with z the FP32 V-cycle output, widened exactly (CG.cpp:110, 113).
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_dot32(int64_t n, const double *x, const float *y,
                                                   double *partials, unsigned *ticket,
                                                   double *scalars, double *out, int finish) {
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = gtid();
  for (; i + 3 * stride < n; i += 4 * stride) {
    double a[4];
    float b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = mload<STREAM>(x + i + u * stride);
      b[u] = mload<STREAM>(y + i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) s += a[u] * (double)b[u];
  }
  for (; i < n; i += stride) s += mload<STREAM>(x + i) * (double)mload<STREAM>(y + i);
  s = block_sum(s);
  grid_finish(s, partials, ticket, scalars, out, finish);
}

// p = z + beta p with z in FP32, widened exactly (CG.cpp:115).
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_update_p32(int64_t n, const float *__restrict__ z,
                                                        double *__restrict__ p,
                                                        const double *__restrict__ scalars) {
  const int64_t i = gtid();
  if (i < n)
    mstore<STREAM>(p + i, (double)mload<STREAM>(z + i) + scalars[kBeta] * mload<STREAM>(p + i));
}

// r = b - Ap and r.r (CG.cpp:88-89)
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_residual0(
    int64_t n, const double *__restrict__ b, const double *__restrict__ Ap,
    double *__restrict__ r, double *partials, unsigned *ticket,
    double *scalars) {
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = gtid();
  for (; i + 3 * stride < n; i += 4 * stride) {
    double bv[4], av[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      bv[u] = mload<STREAM>(b + i + u * stride);
      av[u] = mload<STREAM>(Ap + i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double ri = bv[u] + -1.0 * av[u];
      mstore<STREAM>(r + i + u * stride, ri);
      s += ri * ri;
    }
  }
  for (; i < n; i += stride) {
    const double ri = mload<STREAM>(b + i) + -1.0 * mload<STREAM>(Ap + i);
    mstore<STREAM>(r + i, ri);
    s += ri * ri;
  }
  s = block_sum(s);
  grid_finish(s, partials, ticket, scalars, &scalars[kNormr2], kFinishStore);
}

// r = b - A x, r32 = (float)r and one r.r partial per row chunk in one pass
// (CG.cpp:86-89's residual, as the mixed CG recomputes it in FP64): k_spmv's
// row sums, k_residual0's r and the FP32 rounding, without writing and
// reading A x and r again; k_chunk_dot sums the partials.
template <bool STREAM, bool C16>
__global__ __launch_bounds__(kBlock) void k_spmv_resid(
    const int64_t *__restrict__ chunks, int64_t n, Col16 C, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ x, const double *__restrict__ b,
    double *__restrict__ r, float *__restrict__ r32, double *__restrict__ cpart, bool xcd) {
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, xcd);
  const int64_t i = chunks[2 * blk] + threadIdx.x;
  double s = 0.0;
  if (i < chunks[2 * blk + 1]) {
    const double y = C16 ? row_dot16<STREAM>(C, col, val, x, n, i) : row_dot<STREAM>(col, val, x, n, i);
    const double ri = mload<STREAM>(b + i) + -1.0 * y;
    mstore<STREAM>(r + i, ri);
    mstore<STREAM>(r32 + i, (float)ri);
    s = ri * ri;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) cpart[blk] = s;
}

// k_spmv_resid from the FP16 values and the FP64 diagonal of a matrix whose
// off-diagonal values FP16 holds exactly (row_dot_h, or row_dot_hp on 8-bit
// tiles where vp is non-null): the same products in the same order, so the
// same r.
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_spmv_resid_h(
    const int64_t *__restrict__ chunks, int64_t n, Col16 C, const uint32_t *__restrict__ vc,
    const uint32_t *__restrict__ vp, const uint8_t *__restrict__ dslot,
    const double *__restrict__ diag, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ x, const double *__restrict__ b,
    double *__restrict__ r, float *__restrict__ r32, double *__restrict__ cpart, bool xcd) {
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, xcd);
  const int64_t i = chunks[2 * blk] + threadIdx.x;
  double s = 0.0;
  if (i < chunks[2 * blk + 1]) {
    const double y = vp ? row_dot_hp<STREAM>(C, vp, vc, dslot, diag, col, val, x, n, i)
                        : row_dot_h<STREAM>(C, vc, dslot, diag, col, val, x, n, i);
    const double ri = mload<STREAM>(b + i) + -1.0 * y;
    mstore<STREAM>(r + i, ri);
    mstore<STREAM>(r32 + i, (float)ri);
    s = ri * ri;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) cpart[blk] = s;
}

// p = z + beta p (CG.cpp:115)
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_update_p(
    int64_t n, const double *__restrict__ z, double *__restrict__ p,
    const double *__restrict__ scalars) {
  const int64_t i = gtid();
  if (i < n) mstore<STREAM>(p + i, mload<STREAM>(z + i) + scalars[kBeta] * mload<STREAM>(p + i));
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_fill(double *x, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return;
  k_fill<<<grid_for(n), kBlock, 0, s>>>(x, n, v);
  HPCGX_LAUNCHED();
}

void launch_copy(const double *x, double *y, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  k_copy<<<grid_for(n), kBlock, 0, s>>>(x, y, n);
  HPCGX_LAUNCHED();
}

void launch_gather(const int32_t *idx, const double *src, double *dst,
                   int64_t n, hipStream_t s) {
  if (n <= 0) return;
  k_gather<<<grid_for(n), kBlock, 0, s>>>(idx, src, dst, n);
  HPCGX_LAUNCHED();
}

void launch_spmv(const Problem &P, const Level &L, const double *x, double *y,
                 hipStream_t s) {
  const unsigned g = (unsigned)L.nchunks;
  const Col16 C = col16_of(P, L);
  const bool c16 = use_col16(P, L, 0);
  if (L.stream && c16)
    k_spmv<true, true><<<g, kBlock, 0, s>>>(L.chunks, L.n, C, L.col, L.val, x, y, L.xcd);
  else if (L.stream)
    k_spmv<true, false><<<g, kBlock, 0, s>>>(L.chunks, L.n, C, L.col, L.val, x, y, L.xcd);
  else if (c16)
    k_spmv<false, true><<<g, kBlock, 0, s>>>(L.chunks, L.n, C, L.col, L.val, x, y, L.xcd);
  else
    k_spmv<false, false><<<g, kBlock, 0, s>>>(L.chunks, L.n, C, L.col, L.val, x, y, L.xcd);
  HPCGX_LAUNCHED();
}

void launch_symgs_rows(const Problem &P, const Level &L, int64_t begin, int64_t end,
                       const double *r, double *x, hipStream_t s, double *dpart) {
  if (end <= begin) return;
  if (dpart && !use_col16(P, L, begin))
    throw Error(HPCGX_EINVAL, "colour step with r.z partials needs 16-bit column ids");
  if (use_col16(P, L, begin)) {
    const Col16 C = col16_of(P, L);
    const int64_t rows = end - begin;
    // the x-pair colour steps (default shape only): bit 0 of x_pairs for the
    // steps without r.z, bit 1 for the r.z step
    if (L.finest && (P.x_pairs & (dpart ? 2 : 1)) && P.gs_block == kBlock && P.gs_waves) {
      const unsigned g = (unsigned)((rows + kBlock - 1) / kBlock);
      if (dpart)
        k_gs_rz16<true, kBlock, 6, true><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.col, L.val,
                                                               r, x, L.xcd, dpart);
      else
        k_gs_range16<true, true, kBlock, 7, true><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.col,
                                                                        L.val, r, x, L.xcd);
      HPCGX_LAUNCHED();
      return;
    }
    if (L.finest) {  // tuned workgroup size and register cap
      const bool cap = P.gs_waves != 0;
#define HPCGX_GS(BS)                                                                           \
  do {                                                                                         \
    const unsigned g = (unsigned)((rows + BS - 1) / BS);                                       \
    if (dpart && cap)                                                                          \
      k_gs_rz16<true, BS, 6><<<g, BS, 0, s>>>(begin, end, L.n, C, L.col, L.val, r, x, L.xcd, \
                                              dpart);                                          \
    else if (dpart)                                                                            \
      k_gs_rz16<true, BS, 1><<<g, BS, 0, s>>>(begin, end, L.n, C, L.col, L.val, r, x, L.xcd, \
                                              dpart);                                          \
    else if (cap)                                                                              \
      k_gs_range16<true, true, BS, 7><<<g, BS, 0, s>>>(begin, end, L.n, C, L.col, L.val, r,  \
                                                       x, L.xcd);                              \
    else                                                                                       \
      k_gs_range16<true, true, BS, 1><<<g, BS, 0, s>>>(begin, end, L.n, C, L.col, L.val, r,  \
                                                       x, L.xcd);                              \
  } while (0)
      switch (P.gs_block) {
        case 64: HPCGX_GS(64); break;
        case 128: HPCGX_GS(128); break;
        case 512: HPCGX_GS(512); break;
        default: HPCGX_GS(256); break;
      }
#undef HPCGX_GS
      HPCGX_LAUNCHED();
      return;
    }
    const unsigned g = grid_for(rows);
    if (dpart)
      k_gs_rz16<true, kBlock, 6><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.col, L.val, r, x,
                                                      L.xcd, dpart);
    else if (L.stream && (P.x_pairs & 8))  // bit 3: the coarser levels' steps
      k_gs_range16<true, false, kBlock, 7, true><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.col,
                                                                      L.val, r, x, L.xcd);
    else if (L.stream)
      k_gs_range16<true, false, kBlock, 7><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.col,
                                                                L.val, r, x, L.xcd);
    else
      k_gs_range16<false, false, kBlock, 7><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.col,
                                                                 L.val, r, x, L.xcd);
    HPCGX_LAUNCHED();
    return;
  }
  if (L.pair_rows && L.n % 2 == 0 && begin % 2 == 0) {
    const unsigned g2 = grid_for((end - begin + 1) / 2);
    if (L.finest)
      k_gs_range2<true, true><<<g2, kBlock, 0, s>>>(begin, end, L.n, L.col, L.val, r, x);
    else if (L.stream)
      k_gs_range2<true, false><<<g2, kBlock, 0, s>>>(begin, end, L.n, L.col, L.val, r, x);
    else
      k_gs_range2<false, false><<<g2, kBlock, 0, s>>>(begin, end, L.n, L.col, L.val, r, x);
    HPCGX_LAUNCHED();
    return;
  }
  const unsigned g = grid_for(end - begin);
  if (L.finest)
    k_gs_range<true, true><<<g, kBlock, 0, s>>>(begin, end, L.n, L.col, L.val,
                                                r, x, L.xcd);
  else if (L.stream)
    k_gs_range<true, false><<<g, kBlock, 0, s>>>(begin, end, L.n, L.col, L.val,
                                                 r, x, L.xcd);
  else
    k_gs_range<false, false><<<g, kBlock, 0, s>>>(begin, end, L.n, L.col,
                                                  L.val, r, x, L.xcd);
  HPCGX_LAUNCHED();
}

bool launch_symgs_rows_zero(const Problem &P, const Level &L, int64_t begin, int64_t end,
                            const double *r, double *x, hipStream_t s) {
  // x is read at byte offsets of 32 bits
  if (!L.zmask || !use_col16(P, L, begin) || L.n >= ((int64_t)1 << 29)) return false;
  if (end <= begin) return true;
  const Col16 C = col16_of(P, L);
  const int64_t rows = end - begin;
  const bool live = P.zero_live & (L.finest ? 1 : 2);
  if (L.finest) {  // the finest level: the colour steps' workgroup size (gs_block)
#define HPCGX_GZ(BS, LI)                                                                      \
  k_gs_zero16<true, true, BS, LI><<<(unsigned)((rows + BS - 1) / BS), BS, 0, s>>>(            \
      begin, end, L.n, C, L.zmask, L.col, L.val, L.diag, r, x, L.xcd)
    switch (P.gs_block) {
      case 64: live ? HPCGX_GZ(64, true) : HPCGX_GZ(64, false); break;
      case 128: live ? HPCGX_GZ(128, true) : HPCGX_GZ(128, false); break;
      case 512: live ? HPCGX_GZ(512, true) : HPCGX_GZ(512, false); break;
      default: live ? HPCGX_GZ(256, true) : HPCGX_GZ(256, false); break;
    }
#undef HPCGX_GZ
  } else {
#define HPCGX_GZC(ST, LI)                                                                     \
  k_gs_zero16<ST, false, kBlock, LI><<<grid_for(rows), kBlock, 0, s>>>(                       \
      begin, end, L.n, C, L.zmask, L.col, L.val, L.diag, r, x, L.xcd)
    if (L.stream && live) HPCGX_GZC(true, true);
    else if (L.stream) HPCGX_GZC(true, false);
    else if (live) HPCGX_GZC(false, true);
    else HPCGX_GZC(false, false);
#undef HPCGX_GZC
  }
  HPCGX_LAUNCHED();
  return true;
}

void launch_symgs_list(const Level &L, const int32_t *rows, int64_t count,
                       const double *r, double *x, hipStream_t s) {
  if (count <= 0) return;
  k_gs_list<<<grid_for(count), kBlock, 0, s>>>(rows, count, L.n, L.col, L.val,
                                               r, x);
  HPCGX_LAUNCHED();
}

void launch_restriction(const Level &L, const double *rf, const double *Axf,
                        double *rc, hipStream_t s) {
  k_restrict<<<grid_for(L.nc), kBlock, 0, s>>>(L.nc, L.f2c, rf, Axf, rc);
  HPCGX_LAUNCHED();
}

void launch_residual_restriction(const Problem &P, const Level &L, const double *rf,
                                 const double *xf, double *rc, hipStream_t s) {
  if (L.rperm) {
    const Col16 C = col16_of(P, L);
    const bool c16 = use_col16(P, L, 0);
    const unsigned g = grid_for(L.nc);
    if (L.stream && c16 && L.finest && (P.x_pairs & 1))
      k_resid_restrict_f<true, true, true><<<g, kBlock, 0, s>>>(L.nc, L.ftile, L.rperm, L.n, C,
                                                                L.col, L.val, rf, xf, rc);
    else if (L.stream && c16)
      k_resid_restrict_f<true, true><<<g, kBlock, 0, s>>>(L.nc, L.ftile, L.rperm, L.n, C, L.col,
                                                          L.val, rf, xf, rc);
    else if (L.stream)
      k_resid_restrict_f<true, false><<<g, kBlock, 0, s>>>(L.nc, L.ftile, L.rperm, L.n, C, L.col,
                                                           L.val, rf, xf, rc);
    else if (c16)
      k_resid_restrict_f<false, true><<<g, kBlock, 0, s>>>(L.nc, L.ftile, L.rperm, L.n, C, L.col,
                                                           L.val, rf, xf, rc);
    else
      k_resid_restrict_f<false, false><<<g, kBlock, 0, s>>>(L.nc, L.ftile, L.rperm, L.n, C,
                                                            L.col, L.val, rf, xf, rc);
    HPCGX_LAUNCHED();
    return;
  }
  if (L.stream)
    k_resid_restrict<true><<<grid_for(L.nc), kBlock, 0, s>>>(
        L.nc, L.n, L.f2c, L.col, L.val, rf, xf, rc, L.xcd);
  else
    k_resid_restrict<false><<<grid_for(L.nc), kBlock, 0, s>>>(
        L.nc, L.n, L.f2c, L.col, L.val, rf, xf, rc, L.xcd);
  HPCGX_LAUNCHED();
}

void launch_prolongation(const Level &L, const double *xc, double *xf,
                         hipStream_t s) {
  if (L.rperm) {
    k_prolong_f<<<grid_for(L.nc), kBlock, 0, s>>>(L.nc, L.ftile, L.rperm, xc, xf);
    HPCGX_LAUNCHED();
    return;
  }
  k_prolong<<<grid_for(L.nc), kBlock, 0, s>>>(L.nc, L.f2c, xc, xf);
  HPCGX_LAUNCHED();
}

void launch_waxpby(int64_t n, double alpha, const double *x, double beta,
                   const double *y, double *w, hipStream_t s) {
  if (n <= 0) return;
  const unsigned g = (unsigned)((n + kBlock * kWaxE - 1) / (kBlock * kWaxE));
  const bool st = stream_vec(n);
#define HPCGX_WAX(F)                                                                           \
  if (st)                                                                                      \
    k_waxpby<F, true><<<g, kBlock, 0, s>>>(n, alpha, x, beta, y, w);                           \
  else                                                                                         \
    k_waxpby<F, false><<<g, kBlock, 0, s>>>(n, alpha, x, beta, y, w)
  if (alpha == 1.0) {
    HPCGX_WAX(0);
  } else if (beta == 1.0) {
    HPCGX_WAX(1);
  } else {
    HPCGX_WAX(2);
  }
#undef HPCGX_WAX
  HPCGX_LAUNCHED();
}

void launch_dot_raw(int64_t n, const double *x, const double *y, double *parts,
                    double *partials, unsigned *ticket, double *scalars, double *out, int finish,
                    hipStream_t s) {
  const int64_t np = (n + kDotPer - 1) / kDotPer;
  if (parts && np <= kDotParts) {
    if (np > 0 && stream_vec(n))
      k_dot_rows<true><<<(unsigned)np, kBlock, 0, s>>>(n, x, y, parts);
    else if (np > 0)
      k_dot_rows<false><<<(unsigned)np, kBlock, 0, s>>>(n, x, y, parts);
    if (np > 0) HPCGX_LAUNCHED();
    k_chunk_dot<<<chunk_grid(np), kBlock, 0, s>>>(np, parts, partials, ticket, scalars, finish,
                                                   out);
  } else if (stream_vec(n)) {
    k_dot<true><<<reduce_grid(n), kBlock, 0, s>>>(n, x, y, partials, ticket, scalars, out,
                                                  finish);
  } else {
    k_dot<false><<<reduce_grid(n), kBlock, 0, s>>>(n, x, y, partials, ticket, scalars, out,
                                                   finish);
  }
  HPCGX_LAUNCHED();
}

void launch_zero_u32(unsigned *a, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  k_zero_u32<<<grid_for(n), kBlock, 0, s>>>(a, n);
  HPCGX_LAUNCHED();
}

void launch_dot(Problem &P, int64_t n, const double *x, const double *y,
                double *out, int finish, hipStream_t s) {
  launch_dot_raw(n, x, y, P.dot_parts, P.partials, P.ticket, P.scalars, out, finish, s);
}

void launch_partials_dot(Problem &P, const double *part, int64_t count, int finish,
                         hipStream_t s) {
  k_chunk_dot<<<chunk_grid(count), kBlock, 0, s>>>(count, part, P.partials, P.ticket,
                                                    P.scalars, finish);
  HPCGX_LAUNCHED();
}

namespace {
bool chunked(const Problem &P, const Level &L) {
  return P.spmv_dot_chunked && P.chunk_partials && L.nchunks == P.nchunk_partials;
}
void chunk_dot(Problem &P, const Level &L, hipStream_t s) {
  k_chunk_dot<<<chunk_grid(L.nchunks), kBlock, 0, s>>>(L.nchunks, P.chunk_partials, P.partials,
                                                        P.ticket, P.scalars, kFinishAlpha);
  HPCGX_LAUNCHED();
}
}  // namespace

void launch_spmv_dot(Problem &P, const Level &L, const double *p, double *Ap,
                     hipStream_t s) {
  if (chunked(P, L)) {
    const Col16 C = col16_of(P, L);
    const unsigned G = (unsigned)L.nchunks;
#define HPCGX_SDC(ST, C16)                                                                   \
  k_spmv_dot_c<ST, C16><<<G, kBlock, 0, s>>>(L.chunks, L.n, C, L.col, L.val, p, Ap,           \
                                             P.chunk_partials, L.xcd)
    if (use_col16(P, L, 0) && L.finest && L.stream && (P.x_pairs & 1)) {
      k_spmv_dot_c<true, true, true><<<G, kBlock, 0, s>>>(L.chunks, L.n, C, L.col, L.val, p, Ap,
                                                         P.chunk_partials, L.xcd);
    } else if (use_col16(P, L, 0)) {
      if (L.stream) HPCGX_SDC(true, true); else HPCGX_SDC(false, true);
    } else {
      if (L.stream) HPCGX_SDC(true, false); else HPCGX_SDC(false, false);
    }
#undef HPCGX_SDC
    HPCGX_LAUNCHED();
    chunk_dot(P, L, s);
    return;
  }
  int64_t g = std::min<int64_t>(L.nchunks, std::min(P.spmv_dot_blocks, kMaxPartials));
  g = std::max<int64_t>(8, (g + 7) / 8 * 8);
  const Col16 C = col16_of(P, L);
  const bool c16 = use_col16(P, L, 0);
  const unsigned G = (unsigned)g;
  if (L.stream && c16)
    k_spmv_dot<true, true><<<G, kBlock, 0, s>>>(L.nchunks, L.chunks, L.n, C, L.col, L.val, p,
                                                Ap, P.partials, P.ticket, P.scalars);
  else if (L.stream)
    k_spmv_dot<true, false><<<G, kBlock, 0, s>>>(L.nchunks, L.chunks, L.n, C, L.col, L.val, p,
                                                 Ap, P.partials, P.ticket, P.scalars);
  else if (c16)
    k_spmv_dot<false, true><<<G, kBlock, 0, s>>>(L.nchunks, L.chunks, L.n, C, L.col, L.val, p,
                                                 Ap, P.partials, P.ticket, P.scalars);
  else
    k_spmv_dot<false, false><<<G, kBlock, 0, s>>>(L.nchunks, L.chunks, L.n, C, L.col, L.val,
                                                  p, Ap, P.partials, P.ticket, P.scalars);
  HPCGX_LAUNCHED();
}

void launch_spmv_dot_half(Problem &P, const Level &L, const double *p, double *Ap,
                          hipStream_t s) {
  if (!L.vc || !P.half_exact || !use_col16(P, L, 0)) {
    launch_spmv_dot(P, L, p, Ap, s);
    return;
  }
  const Col16 C = col16_of(P, L);
  if (chunked(P, L)) {
    const unsigned G = (unsigned)L.nchunks;
    const uint32_t *vp =
        P.hc_pairs && L.v16p && P.use_c8 && L.byte_tiles > 0 ? L.v16p : nullptr;
    if (L.stream)
      k_spmv_dot_hc<true><<<G, kBlock, 0, s>>>(L.chunks, L.n, C, L.vc, vp, L.dslot, L.diag,
                                               L.col, L.val, p, Ap, P.chunk_partials, L.xcd);
    else
      k_spmv_dot_hc<false><<<G, kBlock, 0, s>>>(L.chunks, L.n, C, L.vc, vp, L.dslot, L.diag,
                                                L.col, L.val, p, Ap, P.chunk_partials, L.xcd);
    HPCGX_LAUNCHED();
    chunk_dot(P, L, s);
    return;
  }
  int64_t g = std::min<int64_t>(L.nchunks, std::min(P.spmv_dot_blocks, kMaxPartials));
  g = std::max<int64_t>(8, (g + 7) / 8 * 8);
  if (L.stream)
    k_spmv_dot_h<true><<<(unsigned)g, kBlock, 0, s>>>(L.nchunks, L.chunks, L.n, C, L.vc,
                                                      L.dslot, L.diag, L.col, L.val, p, Ap,
                                                      P.partials, P.ticket, P.scalars);
  else
    k_spmv_dot_h<false><<<(unsigned)g, kBlock, 0, s>>>(L.nchunks, L.chunks, L.n, C, L.vc,
                                                       L.dslot, L.diag, L.col, L.val, p, Ap,
                                                       P.partials, P.ticket, P.scalars);
  HPCGX_LAUNCHED();
}

void launch_update_xr(Problem &P, int64_t n, double *x, double *r,
                      const double *p, const double *Ap, hipStream_t s, float *r32) {
  const bool st = stream_vec(n);
  if (P.xr_rows && P.chunk_partials && (int64_t)grid_for(n) <= P.nchunk_partials) {
    const unsigned g = grid_for(n);
#define HPCGX_UXRR(R32, ST, WX)                                                                 \
  k_update_xr_rows<R32, ST, WX><<<g, kBlock, 0, s>>>(n, x, r, p, Ap, r32, P.scalars,            \
                                                     P.chunk_partials)
    if (x) {
      if (r32 && st) HPCGX_UXRR(true, true, true);
      else if (r32) HPCGX_UXRR(true, false, true);
      else if (st) HPCGX_UXRR(false, true, true);
      else HPCGX_UXRR(false, false, true);
    } else {
      if (r32 && st) HPCGX_UXRR(true, true, false);
      else if (r32) HPCGX_UXRR(true, false, false);
      else if (st) HPCGX_UXRR(false, true, false);
      else HPCGX_UXRR(false, false, false);
    }
#undef HPCGX_UXRR
    HPCGX_LAUNCHED();
    k_chunk_dot<<<chunk_grid(g), kBlock, 0, s>>>(g, P.chunk_partials, P.partials, P.ticket,
                                                  P.scalars, kFinishStore, &P.scalars[kNormr2]);
    HPCGX_LAUNCHED();
    return;
  }
  const unsigned g = reduce_grid(n);
#define HPCGX_UXR(R32, ST, WX)                                                                \
  k_update_xr<R32, ST, WX><<<g, kBlock, 0, s>>>(n, x, r, p, Ap, r32, P.partials, P.ticket,    \
                                                P.scalars)
  if (x) {
    if (r32 && st) HPCGX_UXR(true, true, true);
    else if (r32) HPCGX_UXR(true, false, true);
    else if (st) HPCGX_UXR(false, true, true);
    else HPCGX_UXR(false, false, true);
  } else {
    if (r32 && st) HPCGX_UXR(true, true, false);
    else if (r32) HPCGX_UXR(true, false, false);
    else if (st) HPCGX_UXR(false, true, false);
    else HPCGX_UXR(false, false, false);
  }
#undef HPCGX_UXR
  HPCGX_LAUNCHED();
}

void launch_update_xp(int64_t n, double *x, const double *z, double *p,
                      const double *scalars, hipStream_t s) {
  if (stream_vec(n))
    k_update_xp<true><<<grid_for(n), kBlock, 0, s>>>(n, x, z, p, scalars);
  else
    k_update_xp<false><<<grid_for(n), kBlock, 0, s>>>(n, x, z, p, scalars);
  HPCGX_LAUNCHED();
}

void launch_step_x(int64_t n, double *x, const double *p, const double *scalars,
                   hipStream_t s) {
  if (stream_vec(n))
    k_step_x<true><<<grid_for(n), kBlock, 0, s>>>(n, x, p, scalars);
  else
    k_step_x<false><<<grid_for(n), kBlock, 0, s>>>(n, x, p, scalars);
  HPCGX_LAUNCHED();
}

void launch_dot32(Problem &P, int64_t n, const double *x, const float *y, double *out,
                  int finish, hipStream_t s) {
  if (stream_vec(n))
    k_dot32<true><<<reduce_grid(n), kBlock, 0, s>>>(n, x, y, P.partials, P.ticket, P.scalars,
                                                    out, finish);
  else
    k_dot32<false><<<reduce_grid(n), kBlock, 0, s>>>(n, x, y, P.partials, P.ticket, P.scalars,
                                                     out, finish);
  HPCGX_LAUNCHED();
}

void launch_update_xp32(int64_t n, double *x, const float *z, double *p,
                        const double *scalars, hipStream_t s) {
  if (stream_vec(n))
    k_update_xp32<true><<<grid_for(n), kBlock, 0, s>>>(n, x, z, p, scalars);
  else
    k_update_xp32<false><<<grid_for(n), kBlock, 0, s>>>(n, x, z, p, scalars);
  HPCGX_LAUNCHED();
}

void launch_update_p32(int64_t n, const float *z, double *p, const double *scalars,
                       hipStream_t s, bool stream) {
  if (stream && stream_vec(n))
    k_update_p32<true><<<grid_for(n), kBlock, 0, s>>>(n, z, p, scalars);
  else
    k_update_p32<false><<<grid_for(n), kBlock, 0, s>>>(n, z, p, scalars);
  HPCGX_LAUNCHED();
}

void launch_update_p(int64_t n, const double *z, double *p,
                     const double *scalars, hipStream_t s, bool stream) {
  if (stream && stream_vec(n))
    k_update_p<true><<<grid_for(n), kBlock, 0, s>>>(n, z, p, scalars);
  else
    k_update_p<false><<<grid_for(n), kBlock, 0, s>>>(n, z, p, scalars);
  HPCGX_LAUNCHED();
}

bool launch_residual_fused(Problem &P, const Level &L, const double *b, const double *x,
                           double *r, float *r32, hipStream_t s) {
  if (!P.resid_fused || !P.chunk_partials || L.nchunks != P.nchunk_partials) return false;
  const unsigned g = (unsigned)L.nchunks;
  const Col16 C = col16_of(P, L);
  const bool c16 = use_col16(P, L, 0);
  const bool h = P.hc_pairs && L.vc && P.half_exact && c16;
  const uint32_t *vp = h && L.v16p && P.use_c8 && L.byte_tiles > 0 ? L.v16p : nullptr;
#define HPCGX_SR(ST, CC)                                                                       \
  k_spmv_resid<ST, CC><<<g, kBlock, 0, s>>>(L.chunks, L.n, C, L.col, L.val, x, b, r, r32,      \
                                            P.chunk_partials, L.xcd)
#define HPCGX_SRH(ST)                                                                          \
  k_spmv_resid_h<ST><<<g, kBlock, 0, s>>>(L.chunks, L.n, C, L.vc, vp, L.dslot, L.diag, L.col,  \
                                          L.val, x, b, r, r32, P.chunk_partials, L.xcd)
  if (h && L.stream) HPCGX_SRH(true);
  else if (h) HPCGX_SRH(false);
  else if (L.stream && c16) HPCGX_SR(true, true);
  else if (L.stream) HPCGX_SR(true, false);
  else if (c16) HPCGX_SR(false, true);
  else HPCGX_SR(false, false);
#undef HPCGX_SR
#undef HPCGX_SRH
  HPCGX_LAUNCHED();
  k_chunk_dot<<<chunk_grid(L.nchunks), kBlock, 0, s>>>(L.nchunks, P.chunk_partials, P.partials,
                                                       P.ticket, P.scalars, kFinishStore,
                                                       &P.scalars[kNormr2]);
  HPCGX_LAUNCHED();
  return true;
}

void launch_residual0(Problem &P, int64_t n, const double *b, const double *Ap,
                      double *r, hipStream_t s) {
  if (stream_vec(n))
    k_residual0<true><<<reduce_grid(n), kBlock, 0, s>>>(n, b, Ap, r, P.partials, P.ticket,
                                                        P.scalars);
  else
    k_residual0<false><<<reduce_grid(n), kBlock, 0, s>>>(n, b, Ap, r, P.partials, P.ticket,
                                                         P.scalars);
  HPCGX_LAUNCHED();
}

namespace {
template <bool STREAM, bool FINEST, int DIR>
void wave_launch(Problem &P, const Level &L, const double *r, double *x,
                 hipStream_t s) {
  const int32_t *tasks = L.tasks[DIR];
  int64_t ntasks = L.ntasks[DIR];
  const int64_t *tile_ptr = L.tile_ptr;
  int S = L.nslabs, C = L.ncolors, chunk_rows = L.wave_chunk;
  unsigned *done = L.done;
  int *err = P.wave_err;
  int64_t n = L.n;
  const int32_t *col = L.col;
  const double *val = L.val;
  void *args[] = {(void *)&tasks, (void *)&ntasks,     (void *)&tile_ptr,
                  (void *)&S,     (void *)&C,          (void *)&chunk_rows,
                  (void *)&done,  (void *)&err,        (void *)&n,
                  (void *)&col,   (void *)&val,        (void *)&r,
                  (void *)&x};
  const unsigned g = (unsigned)std::min<int64_t>(P.wave_grid, ntasks);
  HPCGX_HIP(hipLaunchCooperativeKernel(
      reinterpret_cast<const void *>(&k_gs_wave<STREAM, FINEST, DIR>), dim3(g),
      dim3(kBlock), args, 0, s));
}

template <bool STREAM, bool FINEST>
void wave_launch(Problem &P, const Level &L, int dir, const double *r,
                 double *x, hipStream_t s) {
  if (dir == 0)
    wave_launch<STREAM, FINEST, 0>(P, L, r, x, s);
  else
    wave_launch<STREAM, FINEST, 1>(P, L, r, x, s);
}

template <bool STREAM, bool FINEST, int DIR>
int wave_blocks_per_cu() {
  int nb = 0;
  HPCGX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, reinterpret_cast<const void *>(&k_gs_wave<STREAM, FINEST, DIR>),
      kBlock, 0));
  return nb;
}
}  // namespace

int wave_grid_size() {
  int dev = 0;
  HPCGX_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HPCGX_HIP(hipGetDeviceProperties(&prop, dev));
  const int nb = std::min({wave_blocks_per_cu<true, true, 0>(),
                           wave_blocks_per_cu<true, true, 1>(),
                           wave_blocks_per_cu<true, false, 0>(),
                           wave_blocks_per_cu<true, false, 1>(),
                           wave_blocks_per_cu<false, false, 0>(),
                           wave_blocks_per_cu<false, false, 1>()});
  if (nb < 1) throw Error(HPCGX_EHIP, "wavefront kernel cannot be resident");
  return nb * prop.multiProcessorCount;
}

void launch_symgs_wave(Problem &P, const Level &L, int dir, const double *r,
                       double *x, hipStream_t s) {
  const int64_t nt = (int64_t)L.ncolors * L.nslabs;
  k_zero_u32<<<grid_for(nt), kBlock, 0, s>>>(L.done, nt);
  HPCGX_LAUNCHED();
  if (L.finest)
    wave_launch<true, true>(P, L, dir, r, x, s);
  else if (L.stream)
    wave_launch<true, false>(P, L, dir, r, x, s);
  else
    wave_launch<false, false>(P, L, dir, r, x, s);
}

void launch_replace_diag(Level &L, const double *d, hipStream_t s) {
  k_replace_diag<<<grid_for(L.n), kBlock, 0, s>>>(L.n, L.dslot, d, L.diag,
                                                  L.val);
  HPCGX_LAUNCHED();
}

}  // namespace hpcgx
