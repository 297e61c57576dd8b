// block.hip -- PCG over s independent right-hand sides at once.
//
// No counterpart in the reference (its CG.cpp:59-143 solves one system).
// The s systems share the matrix: every kernel below walks a row's ELL slots
// once and applies them to all s vectors, so a sweep streams the matrix once
// for all right-hand sides.  Each right-hand side keeps its own recurrence
// (its own alpha, beta and residual norm, CG.cpp:102-124) and gets exactly
// the arithmetic the single-vector path gives it: the FP32 V-cycle on FP16
// values (mixed.hip) as preconditioner, FP64 outer iteration, r = b - A x
// recomputed in FP64 every `recompute` iterations.
//
// Multivectors are [n][s] row-major in the level's device row order: the s
// values of a row are adjacent, so the s lanes that own a row read one
// neighbour's s values with one coalesced access and share its matrix entry.
// s must divide 64 (a wavefront then holds whole rows).
#include <algorithm>
#include <chrono>
#include <cmath>

#include "internal.hpp"

namespace hpcgx {
namespace {

typedef _Float16 half_t;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
__device__ inline int64_t gtid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

template <bool STREAM, class T>
__device__ inline T mload(const T *p) {
  if (STREAM) return __builtin_nontemporal_load(p);
  return *p;
}

// ---- per-right-hand-side reductions -------------------------------------
// Thread t of a workgroup owns right-hand side t % s (the grid stride is a
// multiple of s).  The workgroup's s sums go to partials[block * s + k]; the
// last workgroup (ticket) adds them in a fixed order for each k and applies
// the finishing step to slot k of the scalar arrays (rtz, old rtz, alpha,
// beta, |r|^2: s each).
enum BlockFinish { kBStore = 0, kBBeta = 1, kBAlpha = 2 };

// The tail shared by both layouts: partials[blockIdx.x * s + k] are stored
// (and the workgroup synchronised); the last workgroup reduces and finishes.
// Its thread q sums the partials of column k = q % s from blocks q / s,
// q / s + kBlock / s, ... (four loads in flight), then thread k adds the
// kBlock / s sums of column k in thread order: a fixed order.
__device__ inline void finish_tail_s(int s, double *partials, unsigned *ticket, double *sc,
                                     int finish) {
  __shared__ int is_last;
  __shared__ double red[kBlock];
  if (threadIdx.x == 0) {
    __threadfence();
    is_last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  {
    const int k = threadIdx.x % s;
    const unsigned step = kBlock / s, nb = gridDim.x;
    double t = 0.0;
    unsigned b = threadIdx.x / s;
    for (; b + 3 * step < nb; b += 4 * step) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        v[u] = __hip_atomic_load(&partials[(int64_t)(b + u * step) * s + k], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
      for (int u = 0; u < 4; ++u) t += v[u];
    }
    for (; b < nb; b += step)
      t += __hip_atomic_load(&partials[(int64_t)b * s + k], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    red[threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)s) {
    const int k = threadIdx.x;
    double t = 0.0;
    for (int q = k; q < kBlock; q += s) t += red[q];
    double *rtz = sc, *oldrtz = sc + s, *alpha = sc + 2 * s, *beta = sc + 3 * s,
           *normr2 = sc + 4 * s;
    if (finish == kBStore) {
      normr2[k] = t;
    } else if (finish == kBBeta) {  // CG.cpp:112-114
      oldrtz[k] = rtz[k];
      rtz[k] = t;
      beta[k] = t / oldrtz[k];
    } else {  // CG.cpp:119-120
      alpha[k] = rtz[k] / t;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *ticket = 0u;
}

__device__ inline void block_finish_s(double v, int s, double *partials, unsigned *ticket,
                                      double *sc, int finish) {
  __shared__ double red[kBlock];
  red[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x < (unsigned)s) {  // fixed order: rows of the block ascending
    double t = 0.0;
    for (int q = threadIdx.x; q < kBlock; q += s) t += red[q];
    partials[(int64_t)blockIdx.x * s + threadIdx.x] = t;
  }
  __syncthreads();
  finish_tail_s(s, partials, ticket, sc, finish);
}

// ---- W lanes per row, 16 bytes of right-hand sides per lane ---------------
// For S in {4, 8, 16}: the W = S * (value bytes) / 16 lanes of a row each
// hold V = 16 / (value bytes) of its S right-hand sides (4 FP32 or 2 FP64
// values), so one load instruction of a wavefront reads 64/W whole rows of a
// multivector, one contiguous kilobyte.  The row's ELL slots are read by its
// W lanes at one address (the single-vector kernels' packed FP16 words and
// narrow column ids where the level has them).  A wavefront's rows lie in
// one tile.  Per column the arithmetic is the single-vector kernels'
// (mixed.hip, kernels.hip), term for term.
__device__ inline int tile_of(int64_t i) {
  return __builtin_amdgcn_readfirstlane((int)(i / kTile));
}
__device__ inline half_t vc_val(uint32_t w) {
  return __builtin_bit_cast(half_t, (uint16_t)(w & 0xffffu));
}
template <int S>
__device__ inline void ldv(const float *p, float (&v)[S]) {
#pragma unroll
  for (int q = 0; q < S / 4; ++q) {
    const float4 t = *reinterpret_cast<const float4 *>(p + 4 * q);
    v[4 * q] = t.x;
    v[4 * q + 1] = t.y;
    v[4 * q + 2] = t.z;
    v[4 * q + 3] = t.w;
  }
}
template <int S>
__device__ inline void stv(float *p, const float (&v)[S]) {
#pragma unroll
  for (int q = 0; q < S / 4; ++q)
    *reinterpret_cast<float4 *>(p + 4 * q) = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2],
                                                         v[4 * q + 3]);
}
template <int S>
__device__ inline void ldv(const double *p, double (&v)[S]) {
#pragma unroll
  for (int q = 0; q < S / 2; ++q) {
    const double2 t = *reinterpret_cast<const double2 *>(p + 2 * q);
    v[2 * q] = t.x;
    v[2 * q + 1] = t.y;
  }
}
template <int S>
__device__ inline void stv(double *p, const double (&v)[S]) {
#pragma unroll
  for (int q = 0; q < S / 2; ++q)
    *reinterpret_cast<double2 *>(p + 2 * q) = make_double2(v[2 * q], v[2 * q + 1]);
}

// f(a, c) for every slot of row i in slot order: a the FP16 value (FP32),
// c the column.  C16: packed words of a narrow tile (the caller's wavefront
// lies in one tile), else the FP16 array and 32-bit ids.
template <bool STREAM, bool C16, class F>
__device__ inline void slots32(const Col16 &C, const uint32_t *__restrict__ vc,
                               const int32_t *__restrict__ col,
                               const half_t *__restrict__ val16, int64_t n, int64_t i, F &&f) {
  const int t = C16 ? tile_of(i) : 0;
  if (C16 && C.info[t] != kTile32) {
    const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const uint32_t w = mload<STREAM>(vc + (int64_t)j * n + i);
      f((float)vc_val(w), base[j] + (int32_t)(w >> 16));
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t e = (int64_t)j * n + i;
      f((float)mload<STREAM>(val16 + e), mload<STREAM>(col + e));
    }
  }
}

// Row i, columns [o, o + 4): mixed.hip's gs32_row for each column.  Always
// inlined: with the x-sharing kernels calling it too the compiler made it a
// function of its own, and a call per row took the colour step from 99 to
// 31 registers and from about 190 to 264 us (s = 8, 256^3).
template <bool STREAM, int S, bool C16>
__device__ __attribute__((always_inline)) inline void bgs32_row(int64_t i, int o, int64_t n, const Col16 &C,
                                 const uint32_t *__restrict__ vc,
                                 const int32_t *__restrict__ col,
                                 const half_t *__restrict__ val16,
                                 const float *__restrict__ diag32,
                                 const float *__restrict__ R, float *X) {
  constexpr int V = 4;
  float sum[V], xv[V];
  ldv<V>(R + i * S + o, sum);
  slots32<STREAM, C16>(C, vc, col, val16, n, i, [&](float a, int32_t c) {
    ldv<V>(X + (int64_t)c * S + o, xv);
#pragma unroll
    for (int k = 0; k < V; ++k) sum[k] -= a * xv[k];
  });
  const float d = diag32[i];
#pragma unroll
  for (int k = 0; k < V; ++k) sum[k] = sum[k] / d;
  stv<V>(X + i * S + o, sum);
}

// ---- the first forward sweep of a block V-cycle level from X = 0 ----------
// kernels.hip's zero-x colour step for W lanes per row: the slots that
// Level::zmask leaves out of the tile, and the X rows at or beyond `begin`
// (colours not yet updated in this sweep), are read through buffer views
// whose range check returns 0 without a memory access, so every row's loads
// still go out together and X needs no fill.  Skipped terms are a * 0.
// Needs n * S * 4 < 2^32 (byte offsets into X).
typedef __amdgpu_buffer_rsrc_t BufView;
constexpr int kBufWord3 = 0x00020000;  // buffer descriptor word 3 for gfx9 raw loads
__device__ inline BufView buf_view(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kBufWord3);
}
template <bool STREAM>
__device__ inline uint32_t buf_u32(BufView v, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(v, off, 0, STREAM ? 2 : 0);
}
template <bool STREAM>
__device__ inline uint32_t buf_u16(BufView v, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b16(v, off, 0, STREAM ? 2 : 0);
}
__device__ inline float4 buf_f32x4(BufView v, uint32_t off) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(v, off, 0, 0));
}

template <bool STREAM, int S>
__device__ __attribute__((always_inline)) inline void bgs32_row_zero(
    uint32_t zm, int64_t begin, int64_t i, int o, int64_t n, const Col16 &C,
    const uint32_t *__restrict__ vc, const int32_t *__restrict__ col,
    const half_t *__restrict__ val16, const float *__restrict__ diag32,
    const float *__restrict__ R, float *X) {
  constexpr int V = 4;
  const int t = tile_of(i);
  const int64_t tb = (int64_t)t * kTile;
  const uint32_t row = (uint32_t)(i - tb);
  const uint32_t xlim = (uint32_t)(begin * S * 4);
  float sum[V];
  ldv<V>(R + i * S + o, sum);
  auto term = [&](bool live, float a, int32_t c) {
    const float4 xv = buf_f32x4(buf_view(X, live ? xlim : 0u), (uint32_t)(c * S + o) * 4u);
    sum[0] -= a * xv.x;
    sum[1] -= a * xv.y;
    sum[2] -= a * xv.z;
    sum[3] -= a * xv.w;
  };
  if (C.info[t] != kTile32) {
    const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const bool live = (zm >> j) & 1u;
      const uint32_t w =
          buf_u32<STREAM>(buf_view(vc + (int64_t)j * n + tb, live ? 256u : 0u), row * 4u);
      term(live, (float)vc_val(w), base[j] + (int32_t)(w >> 16));
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const bool live = (zm >> j) & 1u;
      const int32_t c = (int32_t)buf_u32<STREAM>(
          buf_view(col + (int64_t)j * n + tb, live ? 256u : 0u), row * 4u);
      const half_t a = __builtin_bit_cast(
          half_t, (uint16_t)buf_u16<STREAM>(buf_view(val16 + (int64_t)j * n + tb,
                                                     live ? 128u : 0u), row * 2u));
      term(live, (float)a, c);
    }
  }
  const float d = diag32[i];
#pragma unroll
  for (int k = 0; k < V; ++k) sum[k] = sum[k] / d;
  stv<V>(X + i * S + o, sum);
}

// bgs32_row_zero over the live slots only (tune key "zero_live"): Level::
// zmask's set bits in ascending slot order, CH per chunk, the chunk's
// unused entries read through empty views (0 * 0), so the loads of dead
// slots are not issued.  NARROW: packed words (else 32-bit ids), decided
// once per tile outside the loop.
template <bool STREAM, int S, bool NARROW, int CH>
__device__ __attribute__((always_inline)) inline void bzero_live_sum(
    uint32_t zm, int t, int64_t tb, uint32_t row, uint32_t xlim, int o, int64_t n,
    const Col16 &C, const uint32_t *__restrict__ vc, const int32_t *__restrict__ col,
    const half_t *__restrict__ val16, const float *X, float *sum) {
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
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
    // the chunk's index and value words, then its X rows, then the sums
    uint32_t wa[CH], wb[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int j = js[q];
      if (NARROW) {
        wa[q] = buf_u32<STREAM>(buf_view(vc + (int64_t)j * n + tb, ok[q] ? 256u : 0u), row * 4u);
      } else {
        wa[q] = buf_u32<STREAM>(buf_view(col + (int64_t)j * n + tb, ok[q] ? 256u : 0u), row * 4u);
        wb[q] = buf_u16<STREAM>(buf_view(val16 + (int64_t)j * n + tb, ok[q] ? 128u : 0u),
                                row * 2u);
      }
    }
    float a[CH];
    float4 xv[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int j = js[q];
      int32_t c;
      if (NARROW) {
        a[q] = (float)vc_val(wa[q]);
        c = base[j] + (int32_t)(wa[q] >> 16);
      } else {
        c = (int32_t)wa[q];
        a[q] = (float)__builtin_bit_cast(half_t, (uint16_t)wb[q]);
      }
      xv[q] = buf_f32x4(buf_view(X, ok[q] ? xlim : 0u), (uint32_t)(c * S + o) * 4u);
    }
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      sum[0] -= a[q] * xv[q].x;
      sum[1] -= a[q] * xv[q].y;
      sum[2] -= a[q] * xv[q].z;
      sum[3] -= a[q] * xv[q].w;
    }
  }
}

template <bool STREAM, int S, int CH>
__device__ __attribute__((always_inline)) inline void bgs32_row_zero_live(
    uint32_t zm, int64_t begin, int64_t i, int o, int64_t n, const Col16 &C,
    const uint32_t *__restrict__ vc, const int32_t *__restrict__ col,
    const half_t *__restrict__ val16, const float *__restrict__ diag32,
    const float *__restrict__ R, float *X) {
  constexpr int V = 4;
  // a wavefront's 64 / W rows lie in one tile (begin is a multiple of
  // kTile), so the tile and its mask are uniform: the slot loop runs on
  // scalar registers
  const int t = __builtin_amdgcn_readfirstlane(tile_of(i));
  zm = __builtin_amdgcn_readfirstlane(zm);
  const int64_t tb = (int64_t)t * kTile;
  const uint32_t row = (uint32_t)(i - tb);
  const uint32_t xlim = (uint32_t)(begin * S * 4);
  float sum[V];
  ldv<V>(R + i * S + o, sum);
  if (C.info[t] != kTile32)
    bzero_live_sum<STREAM, S, true, CH>(zm, t, tb, row, xlim, o, n, C, vc, col, val16, X, sum);
  else
    bzero_live_sum<STREAM, S, false, CH>(zm, t, tb, row, xlim, o, n, C, vc, col, val16, X, sum);
  const float d = diag32[i];
#pragma unroll
  for (int k = 0; k < V; ++k) sum[k] = sum[k] / d;
  stv<V>(X + i * S + o, sum);
}

// A colour step of the first forward sweep from X = 0, rows [begin, end)
// (begin a multiple of kTile).
// LIVE: the live slots only (bgs32_row_zero_live), else every slot (bgs32_row_zero)
template <bool STREAM, int S, bool LIVE = false>
__global__ __launch_bounds__(kBlock) void k_bgs32_rows_zero(
    int64_t begin, int64_t end, int64_t n, Col16 C, const uint32_t *__restrict__ zmask,
    const uint32_t *__restrict__ vc, const int32_t *__restrict__ col,
    const half_t *__restrict__ val16, const float *__restrict__ diag32,
    const float *__restrict__ R, float *X) {
  constexpr int W = S / 4;
  const int64_t g = gtid();
  const int64_t i = begin + g / W;
  if (i >= end) return;
  if (LIVE)
    bgs32_row_zero_live<STREAM, S, 8>(zmask[tile_of(i)], begin, i, (int)(g % W) * 4, n, C, vc,
                                      col, val16, diag32, R, X);
  else
    bgs32_row_zero<STREAM, S>(zmask[tile_of(i)], begin, i, (int)(g % W) * 4, n, C, vc, col,
                              val16, diag32, R, X);
}

// Whether tile t has all its offset words affine (slot j's column is its
// base plus the row's position in the tile) and slot 3m + 2's base one
// above slot 3m's for every m: on the 27-point grid the x - 1 and x + 1
// neighbours, consecutive rows of one colour block (uniform per tile).
__device__ inline bool bpairs_tile(const Col16 &C, int t) {
  if (C.info[t] != kTile8 || !C.amask || C.amask[t] != (1u << kSlotWords) - 1u) return false;
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
  bool ok = true;
#pragma unroll
  for (int m = 0; m < kSlots / 3; ++m) ok = ok && base[3 * m + 2] == base[3 * m] + 1;
  return ok;
}

// bgs32_row on a wavefront whose 64 / W rows are all present and whose tile
// passes bpairs_tile: slot 3m + 2's x rows are slot 3m's one row (W lanes)
// further on, so they come from slot 3m's values moved down W lanes
// (ds_bpermute); the wavefront's last row reads its own.  Same terms in the
// same order.  The loop over the nine slot triples stays rolled: unrolled,
// the compiler hoists every triple's loads above the lane moves and needs
// some 180 registers.
template <bool STREAM, int S>
__device__ __attribute__((always_inline)) inline void bgs32_row_pairs(int64_t i, int o, int64_t n, const Col16 &C,
                                       const uint32_t *__restrict__ vc,
                                       const float *__restrict__ diag32,
                                       const float *__restrict__ R, float *X) {
  constexpr int V = 4, W = S / 4, ROWS = 64 / W;
  const int t = tile_of(i);
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
  const int lane = (int)(threadIdx.x & 63);
  const bool last = lane >= 64 - W;
  const int64_t i0 = i - lane / W;  // the wavefront's first row
  const int32_t plast = (int32_t)(__builtin_amdgcn_readfirstlane((int)(i0 % kTile)) + ROWS - 1);
  float sum[V];
  ldv<V>(R + i * S + o, sum);
#pragma unroll 1
  for (int m = 0; m < kSlots / 3; ++m) {
    const uint32_t w0 = mload<STREAM>(vc + (int64_t)(3 * m) * n + i);
    const uint32_t w1 = mload<STREAM>(vc + (int64_t)(3 * m + 1) * n + i);
    const uint32_t w2 = mload<STREAM>(vc + (int64_t)(3 * m + 2) * n + i);
    const int32_t c0 = base[3 * m] + (int32_t)(w0 >> 16);
    const int32_t c1 = base[3 * m + 1] + (int32_t)(w1 >> 16);
    float x0[V], x1[V], xl[V];
    ldv<V>(X + (int64_t)c0 * S + o, x0);
    ldv<V>(X + (int64_t)c1 * S + o, x1);
    ldv<V>(X + (int64_t)(base[3 * m + 2] + plast) * S + o, xl);
    const float a0 = (float)vc_val(w0), a1 = (float)vc_val(w1), a2 = (float)vc_val(w2);
#pragma unroll
    for (int k = 0; k < V; ++k) sum[k] -= a0 * x0[k];
#pragma unroll
    for (int k = 0; k < V; ++k) sum[k] -= a1 * x1[k];
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const float up = __shfl_down(x0[k], W, 64);
      sum[k] -= a2 * (last ? xl[k] : up);
    }
  }
  const float d = diag32[i];
#pragma unroll
  for (int k = 0; k < V; ++k) sum[k] = sum[k] / d;
  stv<V>(X + i * S + o, sum);
}

// One colour step, rows [begin, end) (begin a multiple of kTile when C16).
// PAIRS: bgs32_row_pairs on the wavefronts where it applies.
template <bool STREAM, int S, bool C16, bool PAIRS = false>
__global__ __launch_bounds__(kBlock) void k_bgs32_rows(int64_t begin, int64_t end, int64_t n,
                                                       Col16 C, const uint32_t *__restrict__ vc,
                                                       const int32_t *__restrict__ col,
                                                       const half_t *__restrict__ val16,
                                                       const float *__restrict__ diag32,
                                                       const float *__restrict__ R, float *X) {
  constexpr int W = S / 4;
  const int64_t g = gtid();
  const int64_t i = begin + g / W;
  if (PAIRS) {
    // the wavefront's first row; all its 64 / W rows present (uniform)
    const int64_t i0 = begin + ((int64_t)blockIdx.x * kBlock + (threadIdx.x & ~63u)) / W;
    if (i0 + 64 / W <= end && bpairs_tile(C, tile_of(i0))) {
      bgs32_row_pairs<STREAM, S>(i, (int)(g % W) * 4, n, C, vc, diag32, R, X);
      return;
    }
  }
  if (i >= end) return;
  bgs32_row<STREAM, S, C16>(i, (int)(g % W) * 4, n, C, vc, col, val16, diag32, R, X);
}

// rc = rf[f2c] - (A xf)[f2c] in fine order (Level::ftile, rperm): mixed.hip's
// row_dot32 for each column.
template <bool STREAM, int S, bool C16>
__global__ __launch_bounds__(kBlock) void k_bresid_restrict32_rows(
    int64_t nc, const int32_t *__restrict__ ftile, const int32_t *__restrict__ rperm, int64_t n,
    Col16 C, const uint32_t *__restrict__ vc, const int32_t *__restrict__ col,
    const half_t *__restrict__ val16, const float *__restrict__ diag32,
    const float *__restrict__ RF, const float *__restrict__ XF, float *__restrict__ RC) {
  constexpr int V = 4, W = S / V;
  const int64_t g = gtid();
  const int64_t t = g / W;
  const int o = (int)(g % W) * V;
  if (t >= nc) return;
  const int64_t f = fine_row(ftile, t);
  float a[V], xv[V];
#pragma unroll
  for (int k = 0; k < V; ++k) a[k] = 0.0f;
  slots32<STREAM, C16>(C, vc, col, val16, n, f, [&](float v, int32_t c) {
    ldv<V>(XF + (int64_t)c * S + o, xv);
#pragma unroll
    for (int k = 0; k < V; ++k) a[k] += v * xv[k];
  });
  const float d = diag32[f];
  ldv<V>(XF + f * S + o, xv);
#pragma unroll
  for (int k = 0; k < V; ++k) a[k] += d * xv[k];
  float rf[V];
  ldv<V>(RF + f * S + o, rf);
#pragma unroll
  for (int k = 0; k < V; ++k) rf[k] = rf[k] - a[k];
  stv<V>(RC + (int64_t)rperm[t] * S + o, rf);
}

// xf[f2c] += xc in fine order
template <int S>
__global__ __launch_bounds__(kBlock) void k_bprolong32_rows(int64_t nc,
                                                            const int32_t *__restrict__ ftile,
                                                            const int32_t *__restrict__ rperm,
                                                            const float *__restrict__ XC,
                                                            float *__restrict__ XF) {
  constexpr int V = 4, W = S / V;
  const int64_t g = gtid();
  const int64_t t = g / W;
  const int o = (int)(g % W) * V;
  if (t >= nc) return;
  const int64_t f = fine_row(ftile, t);
  float a[V], b[V];
  ldv<V>(XF + f * S + o, a);
  ldv<V>(XC + (int64_t)rperm[t] * S + o, b);
#pragma unroll
  for (int k = 0; k < V; ++k) a[k] += b[k];
  stv<V>(XF + f * S + o, a);
}

// Y = A X in FP64 (ComputeSPMV_ref.cpp:62-69's order for each column) and,
// when `dot`, p_k . (A p)_k and alpha_k (CG.cpp:118-120).  H: the FP16
// off-diagonal values and the FP64 diagonal (exactly the FP64 values when
// the off-diagonal values are exact in FP16), else the FP64 ELL.  Work-items
// gtid(), gtid() + stride, ... (stride a multiple of kTile), W per row.
template <bool STREAM, int S, bool H>
__device__ inline void bspmv_row(int64_t i, int o, int64_t n, const Col16 &C,
                                 const uint32_t *__restrict__ vc,
                                 const uint8_t *__restrict__ dslot,
                                 const double *__restrict__ diag,
                                 const int32_t *__restrict__ col,
                                 const double *__restrict__ val,
                                 const double *__restrict__ X, double (&y)[2]) {
  constexpr int V = 2;
  double xv[V];
#pragma unroll
  for (int k = 0; k < V; ++k) y[k] = 0.0;
  const int t = H ? tile_of(i) : 0;
  if (H && C.info[t] != kTile32) {
    const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
    const int ds = dslot[i];
    const double di = diag[i];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const uint32_t w = mload<STREAM>(vc + (int64_t)j * n + i);
      const int32_t c = base[j] + (int32_t)(w >> 16);
      const double a = j == ds ? di : (double)vc_val(w);
      ldv<V>(X + (int64_t)c * S + o, xv);
#pragma unroll
      for (int k = 0; k < V; ++k) y[k] += a * xv[k];
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t e = (int64_t)j * n + i;
      const double a = mload<STREAM>(val + e);
      ldv<V>(X + (int64_t)mload<STREAM>(col + e) * S + o, xv);
#pragma unroll
      for (int k = 0; k < V; ++k) y[k] += a * xv[k];
    }
  }
}

// Column sums of the workgroup's acc values (thread q holds columns
// (q % W) * V + v): for column k, the threads holding it in thread order.
// Valid in threads k < S.
template <int S>
__device__ inline double bcolumn_sum(const double (&acc)[2]) {
  constexpr int V = 2, W = S / V;
  __shared__ double red[kBlock][V];
#pragma unroll
  for (int k = 0; k < V; ++k) red[threadIdx.x][k] = acc[k];
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x < (unsigned)S) {
    const int k = threadIdx.x, sub = k / V, v = k % V;
    for (int q = sub; q < kBlock; q += W) t += red[q][v];
  }
  __syncthreads();
  return t;
}

template <bool STREAM, int S, bool H>
__global__ __launch_bounds__(kBlock) void k_bspmv_rows(
    int64_t n, Col16 C, const uint32_t *__restrict__ vc, const uint8_t *__restrict__ dslot,
    const double *__restrict__ diag, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ X, double *__restrict__ Y,
    bool dot, double *partials, unsigned *ticket, double *sc) {
  constexpr int V = 2, W = S / V;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double acc[V];
#pragma unroll
  for (int k = 0; k < V; ++k) acc[k] = 0.0;
  const int o = (int)(threadIdx.x % W) * V;  // stride and blockDim are multiples of W
  for (int64_t g = gtid(); g < n * W; g += stride) {
    const int64_t i = g / W;
    double y[V], xv[V];
    bspmv_row<STREAM, S, H>(i, o, n, C, vc, dslot, diag, col, val, X, y);
    stv<V>(Y + i * S + o, y);
    if (dot) {
      ldv<V>(X + i * S + o, xv);
#pragma unroll
      for (int k = 0; k < V; ++k) acc[k] += xv[k] * y[k];
    }
  }
  if (!dot) return;
  const double t = bcolumn_sum<S>(acc);
  if (threadIdx.x < (unsigned)S) partials[(int64_t)blockIdx.x * S + threadIdx.x] = t;
  __syncthreads();
  finish_tail_s(S, partials, ticket, sc, kBAlpha);
}

// The same with the rows of one chunk (Level::chunks) per W workgroups, in
// the chunk list's order: workgroup q takes rows [u * kBlock / W, ...) of
// chunk q / W, u = q % W.  With dot, its column sums go to cpart[q * S + k]
// for k_chunk_bdot.
template <bool STREAM, int S, bool H>
__global__ __launch_bounds__(kBlock) void k_bspmv_chunks(
    const int64_t *__restrict__ chunks, int64_t n, Col16 C, const uint32_t *__restrict__ vc,
    const uint8_t *__restrict__ dslot, const double *__restrict__ diag,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ X, double *__restrict__ Y, bool dot, double *__restrict__ cpart) {
  constexpr int V = 2, W = S / V;
  const int64_t q = blockIdx.x, b = q / W;
  const int64_t i = chunks[2 * b] + (q % W) * (kBlock / W) + threadIdx.x / W;
  const int o = (int)(threadIdx.x % W) * V;
  double acc[V] = {0.0, 0.0};
  if (i < chunks[2 * b + 1]) {
    double y[V], xv[V];
    bspmv_row<STREAM, S, H>(i, o, n, C, vc, dslot, diag, col, val, X, y);
    stv<V>(Y + i * S + o, y);
    if (dot) {
      ldv<V>(X + i * S + o, xv);
#pragma unroll
      for (int k = 0; k < V; ++k) acc[k] = xv[k] * y[k];
    }
  }
  if (!dot) return;
  const double t = bcolumn_sum<S>(acc);
  if (threadIdx.x < (unsigned)S) cpart[q * S + threadIdx.x] = t;
}

// R = B - A X, R32 = (float)R and the column sums of R.R over the rows of
// one chunk per W workgroups (k_bspmv_chunks's rows; CG.cpp:86-89 per
// column, as the periodic recompute needs it): one pass, where the SpMM and
// k_bresidual0 wrote and read A X again.  Column sums to cpart[q * S + k].
template <bool STREAM, int S, bool H>
__global__ __launch_bounds__(kBlock) void k_bspmv_resid_chunks(
    const int64_t *__restrict__ chunks, int64_t n, Col16 C, const uint32_t *__restrict__ vc,
    const uint8_t *__restrict__ dslot, const double *__restrict__ diag,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ X, const double *__restrict__ Bv, double *__restrict__ R,
    float *__restrict__ R32, double *__restrict__ cpart) {
  constexpr int V = 2, W = S / V;
  const int64_t q = blockIdx.x, b = q / W;
  const int64_t i = chunks[2 * b] + (q % W) * (kBlock / W) + threadIdx.x / W;
  const int o = (int)(threadIdx.x % W) * V;
  double acc[V] = {0.0, 0.0};
  if (i < chunks[2 * b + 1]) {
    double y[V], bv[V];
    bspmv_row<STREAM, S, H>(i, o, n, C, vc, dslot, diag, col, val, X, y);
    ldv<V>(Bv + i * S + o, bv);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const double r = bv[k] - y[k];
      R[i * S + o + k] = r;
      R32[i * S + o + k] = (float)r;
      acc[k] = r * r;
    }
  }
  const double t = bcolumn_sum<S>(acc);
  if (threadIdx.x < (unsigned)S) cpart[q * S + threadIdx.x] = t;
}

// p_k . (A p)_k from k_bspmv_chunks's partials (a launch of its own, so
// their stores are ordered before these loads), in fixed order: thread t of
// workgroup w sums column t % s of parts lo + t / s, + kBlock / s, ...;
// block_finish_s then adds the threads of each column in order.
__global__ __launch_bounds__(kBlock) void k_chunk_bdot(int64_t nparts, int s,
                                                       const double *__restrict__ cpart,
                                                       double *partials, unsigned *ticket,
                                                       double *sc, int finish) {
  const int64_t per = (nparts + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < nparts ? lo + per : nparts;
  const int k = threadIdx.x % s, step = kBlock / s;
  double t = 0.0;
  for (int64_t p = lo + threadIdx.x / s; p < hi; p += step) t += cpart[p * s + k];
  block_finish_s(t, s, partials, ticket, sc, finish);
}

// ---- FP64 outer kernels ----------------------------------------------------
// Y = A X and p_k . (A p)_k (CG.cpp:118-120), slots in the reference's order.
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_bspmv_dot(int64_t n, int s,
                                                      const int32_t *__restrict__ col,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ X,
                                                      double *__restrict__ Y, bool dot,
                                                      double *partials, unsigned *ticket,
                                                      double *sc) {
  const int64_t total = n * s, stride = (int64_t)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (int64_t t = gtid(); t < total; t += stride) {
    const int64_t i = t / s;
    const int k = (int)(t - i * s);
    double sum = 0.0;
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t e = (int64_t)j * n + i;
      sum += mload<STREAM>(val + e) * X[(int64_t)mload<STREAM>(col + e) * s + k];
    }
    Y[t] = sum;
    acc += X[t] * sum;
  }
  if (dot) block_finish_s(acc, s, partials, ticket, sc, kBAlpha);
}

// r = b - Ap, r32 = (float)r, |r|^2 per right-hand side (CG.cpp:88-89)
__global__ __launch_bounds__(kBlock) void k_bresidual0(int64_t total, int s,
                                                       const double *__restrict__ B,
                                                       const double *__restrict__ AP,
                                                       double *__restrict__ R,
                                                       float *__restrict__ R32, double *partials,
                                                       unsigned *ticket, double *sc) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (int64_t t = gtid(); t < total; t += stride) {
    const double r = B[t] - AP[t];
    R[t] = r;
    R32[t] = (float)r;
    acc += r * r;
  }
  block_finish_s(acc, s, partials, ticket, sc, kBStore);
}

// r_k . z_k with z the FP32 V-cycle output widened exactly; beta (CG.cpp:110-114)
__global__ __launch_bounds__(kBlock) void k_bdot32(int64_t total, int s,
                                                   const double *__restrict__ R,
                                                   const float *__restrict__ Z32,
                                                   double *partials, unsigned *ticket,
                                                   double *sc) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (int64_t t = gtid(); t < total; t += stride) acc += R[t] * (double)Z32[t];
  block_finish_s(acc, s, partials, ticket, sc, kBBeta);
}

// p = z (first iteration) or p = z + beta_k p (CG.cpp:109, 115)
__global__ __launch_bounds__(kBlock) void k_bupdate_p32(int64_t total, int s,
                                                        const float *__restrict__ Z32,
                                                        double *__restrict__ Pv,
                                                        const double *__restrict__ sc,
                                                        bool first) {
  const int64_t t = gtid();
  if (t >= total) return;
  const int k = (int)(t % s);
  Pv[t] = first ? (double)Z32[t] : (double)Z32[t] + sc[3 * s + k] * Pv[t];
}

// elements per thread of the block vector kernels that take the workgroup-
// per-kBlock*kBE layout (k_bupdate_xp32, k_bdot32_rows, k_bupdate_xr_rows)
constexpr int kBE = 4;

// Multivector loads and stores of those kernels; NT: non-temporal (the
// multivectors are far larger than the last-level cache)
template <bool NT, typename T>
__device__ inline T vld(const T *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT, typename T>
__device__ inline void vst(T *p, T v) {
  if (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// p = z + beta_k p with the previous iteration's x += alpha_k p taken in
// the same pass (CG.cpp:115, :121; "defer_x"): alpha is still the previous
// iteration's when this runs, before the SpMV's p.Ap sets the new one.
// Taken with the kBE-per-thread layout of k_bupdate_xr_rows (s divides 64,
// so the column is threadIdx.x % s).
template <bool NT>
__global__ __launch_bounds__(kBlock) void k_bupdate_xp32(int64_t total, int s,
                                                         double *__restrict__ X,
                                                         const float *__restrict__ Z32,
                                                         double *__restrict__ Pv,
                                                         const double *__restrict__ sc) {
  const int64_t t0 = (int64_t)blockIdx.x * (kBlock * kBE) + threadIdx.x;
  const int k = (int)(threadIdx.x % s);
  const double alpha = sc[2 * s + k], beta = sc[3 * s + k];
  double xv[kBE], pv[kBE];
  float zv[kBE];
#pragma unroll
  for (int e = 0; e < kBE; ++e) {
    const int64_t t = t0 + (int64_t)e * kBlock;
    xv[e] = t < total ? vld<NT>(X + t) : 0.0;
    pv[e] = t < total ? vld<NT>(Pv + t) : 0.0;
    zv[e] = t < total ? vld<NT>(Z32 + t) : 0.0f;
  }
#pragma unroll
  for (int e = 0; e < kBE; ++e) {
    const int64_t t = t0 + (int64_t)e * kBlock;
    if (t < total) {
      vst<NT>(X + t, xv[e] + alpha * pv[e]);
      vst<NT>(Pv + t, (double)zv[e] + beta * pv[e]);
    }
  }
}

// x += alpha_k p alone: a deferred step before the residual recompute and
// after the last iteration
__global__ __launch_bounds__(kBlock) void k_bstep_x(int64_t total, int s, double *__restrict__ X,
                                                    const double *__restrict__ Pv,
                                                    const double *__restrict__ sc) {
  const int64_t t = gtid();
  if (t >= total) return;
  X[t] = X[t] + sc[2 * s + (int)(t % s)] * Pv[t];
}

// x += alpha_k p; r -= alpha_k Ap; r32 = (float)r; |r|^2 (CG.cpp:121-124)
__global__ __launch_bounds__(kBlock) void k_bupdate_xr(int64_t total, int s,
                                                       double *__restrict__ X,
                                                       double *__restrict__ R,
                                                       const double *__restrict__ Pv,
                                                       const double *__restrict__ AP,
                                                       float *__restrict__ R32, double *partials,
                                                       unsigned *ticket, double *sc) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (int64_t t = gtid(); t < total; t += stride) {
    const int k = (int)(t % s);
    const double alpha = sc[2 * s + k];
    X[t] = X[t] + alpha * Pv[t];
    const double r = R[t] + (-alpha) * AP[t];
    R[t] = r;
    R32[t] = (float)r;
    acc += r * r;
  }
  block_finish_s(acc, s, partials, ticket, sc, kBStore);
}

// ---- the vector kernels, kBE elements per thread ("xr_rows") -------------
// Workgroup w takes elements [w * kBlock * kBE, ...) of the [n][s]
// multivectors, thread q the elements w * kBlock * kBE + q + e * kBlock; its
// column is q % s (s divides 64).  Each workgroup stores its column sums to
// cpart[w * s + k] with a plain store, and k_chunk_bdot, a launch of its
// own, sums them in a fixed order and finishes.  The grid-stride forms above
// keep a small grid and the last workgroup's finish; these launch a
// workgroup per kBlock * kBE elements, so more loads are in flight.

// Column sums of a workgroup in a fixed order: within each wavefront a
// shuffle tree over the lanes of a column (lanes k, k + s, ...), then the
// wavefronts in order.  Valid in threads k < s.
__device__ inline double bwave_column_sum(double v, int s) {
  __shared__ double red[kBlock / 64][64];
  for (int off = 32; off >= s; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < s) red[w][lane] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x < (unsigned)s) {
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) t += red[q][threadIdx.x];
  }
  return t;
}

// r_k . z_k per workgroup (k_bdot32's sums)
template <bool NT>
__global__ __launch_bounds__(kBlock) void k_bdot32_rows(int64_t total, int s,
                                                        const double *__restrict__ R,
                                                        const float *__restrict__ Z32,
                                                        double *__restrict__ cpart) {
  const int64_t t0 = (int64_t)blockIdx.x * (kBlock * kBE) + threadIdx.x;
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < kBE; ++e) {
    const int64_t t = t0 + (int64_t)e * kBlock;
    if (t < total) acc += vld<NT>(R + t) * (double)vld<NT>(Z32 + t);
  }
  const double c = bwave_column_sum(acc, s);
  if (threadIdx.x < (unsigned)s) cpart[(int64_t)blockIdx.x * s + threadIdx.x] = c;
}

// k_bupdate_xr's update, |r|^2 per workgroup; WITH_X false: r, r32 and
// |r|^2 only (x's step deferred to k_bupdate_xp32)
template <bool WITH_X, bool NT>
__global__ __launch_bounds__(kBlock) void k_bupdate_xr_rows(int64_t total, int s,
                                                            double *__restrict__ X,
                                                            double *__restrict__ R,
                                                            const double *__restrict__ Pv,
                                                            const double *__restrict__ AP,
                                                            float *__restrict__ R32,
                                                            const double *__restrict__ sc,
                                                            double *__restrict__ cpart) {
  const int64_t t0 = (int64_t)blockIdx.x * (kBlock * kBE) + threadIdx.x;
  const double alpha = sc[2 * s + (int)(threadIdx.x % s)];
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < kBE; ++e) {
    const int64_t t = t0 + (int64_t)e * kBlock;
    if (t < total) {
      if (WITH_X) vst<NT>(X + t, vld<NT>(X + t) + alpha * vld<NT>(Pv + t));
      const double r = vld<NT>(R + t) + (-alpha) * vld<NT>(AP + t);
      vst<NT>(R + t, r);
      vst<NT>(R32 + t, (float)r);
      acc += r * r;
    }
  }
  const double c = bwave_column_sum(acc, s);
  if (threadIdx.x < (unsigned)s) cpart[(int64_t)blockIdx.x * s + threadIdx.x] = c;
}

// ---- FP32 V-cycle kernels (mixed.hip's arithmetic, s vectors) ------------
__global__ __launch_bounds__(kBlock) void k_bfill32(float *x, int64_t total) {
  const int64_t t = gtid();
  if (t < total) x[t] = 0.0f;
}

// One colour step: rows [begin, end), all s right-hand sides.
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_bgs32(int64_t begin, int64_t end, int64_t n, int s,
                                                  const int32_t *__restrict__ col,
                                                  const half_t *__restrict__ val16,
                                                  const float *__restrict__ diag32,
                                                  const float *__restrict__ R, float *X) {
  const int64_t t = gtid();
  const int64_t i = begin + t / s;
  if (i >= end) return;
  const int k = (int)(t % s);
  float sum = R[i * s + k];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t e = (int64_t)j * n + i;
    sum -= (float)mload<STREAM>(val16 + e) * X[(int64_t)mload<STREAM>(col + e) * s + k];
  }
  X[i * s + k] = sum / diag32[i];
}

// One level of the natural-order schedule (rows listed), all s vectors.
__global__ __launch_bounds__(kBlock) void k_bgs32_list(const int32_t *__restrict__ rows,
                                                       int64_t count, int64_t n, int s,
                                                       const int32_t *__restrict__ col,
                                                       const half_t *__restrict__ val16,
                                                       const float *__restrict__ diag32,
                                                       const float *__restrict__ R, float *X) {
  const int64_t t = gtid();
  if (t >= count * s) return;
  const int64_t i = rows[t / s];
  const int k = (int)(t % s);
  float sum = R[i * s + k];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t e = (int64_t)j * n + i;
    sum -= (float)val16[e] * X[(int64_t)col[e] * s + k];
  }
  X[i * s + k] = sum / diag32[i];
}

// rc = rf[f2c] - (A xf)[f2c] in FP32 (mixed.hip's k_resid_restrict32)
template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_bresid_restrict32(
    int64_t nc, int64_t n, int s, const int32_t *__restrict__ f2c,
    const int32_t *__restrict__ col, const half_t *__restrict__ val16,
    const float *__restrict__ diag32, const float *__restrict__ RF, const float *__restrict__ XF,
    float *__restrict__ RC) {
  const int64_t t = gtid();
  if (t >= nc * s) return;
  const int64_t c = t / s;
  const int k = (int)(t % s);
  const int64_t f = f2c[c];
  float a = 0.0f;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t e = (int64_t)j * n + f;
    a += (float)mload<STREAM>(val16 + e) * XF[(int64_t)mload<STREAM>(col + e) * s + k];
  }
  a += diag32[f] * XF[f * s + k];
  RC[t] = RF[f * s + k] - a;
}

__global__ __launch_bounds__(kBlock) void k_bprolong32(int64_t nc, int s,
                                                       const int32_t *__restrict__ f2c,
                                                       const float *__restrict__ XC,
                                                       float *__restrict__ XF) {
  const int64_t t = gtid();
  if (t >= nc * s) return;
  const int64_t c = t / s;
  const int k = (int)(t % s);
  XF[(int64_t)f2c[c] * s + k] += XC[t];
}

unsigned block_reduce_grid(int64_t total, int s) {
  // whole workgroups of kBlock threads; the stride kBlock*G is a multiple of s
  const int64_t g = std::min<int64_t>((total + kBlock - 1) / kBlock, kReduceBlocks);
  return (unsigned)std::max<int64_t>(1, g);
}

// ---- launchers of the W-lanes-per-row kernels (S in {4, 8, 16}) ---------
bool rows_kernels(int s) { return s == 4 || s == 8 || s == 16; }

template <int S>
void gs_rows_s(const Problem &P, const Level &L, int64_t b, int64_t e, const float *R, float *X,
               hipStream_t st) {
  const Col16 C = col16_of(P, L);
  const bool c16 = L.vc && use_col16(P, L, b);
  const unsigned g = grid_for((e - b) * (S / 4));
  // the shared neighbour rows: 16.22 -> 15.63 ms per block iteration at s = 8
  // (256^3), 30.05 -> 30.79 ms at s = 16 (same job), so s = 8 only
  if (S == 8 && c16 && P.block_pairs && L.finest && L.stream)
    k_bgs32_rows<true, S, true, true><<<g, kBlock, 0, st>>>(b, e, L.n, C, L.vc, L.col, L.val16,
                                                            L.diag32, R, X);
  else if (S == 8 && c16 && P.block_pairs && L.finest)
    k_bgs32_rows<false, S, true, true><<<g, kBlock, 0, st>>>(b, e, L.n, C, L.vc, L.col,
                                                             L.val16, L.diag32, R, X);
  else if (L.stream && c16)
    k_bgs32_rows<true, S, true><<<g, kBlock, 0, st>>>(b, e, L.n, C, L.vc, L.col, L.val16,
                                                      L.diag32, R, X);
  else if (L.stream)
    k_bgs32_rows<true, S, false><<<g, kBlock, 0, st>>>(b, e, L.n, C, L.vc, L.col, L.val16,
                                                       L.diag32, R, X);
  else if (c16)
    k_bgs32_rows<false, S, true><<<g, kBlock, 0, st>>>(b, e, L.n, C, L.vc, L.col, L.val16,
                                                       L.diag32, R, X);
  else
    k_bgs32_rows<false, S, false><<<g, kBlock, 0, st>>>(b, e, L.n, C, L.vc, L.col, L.val16,
                                                        L.diag32, R, X);
  HPCGX_LAUNCHED();
}

template <int S>
void gs_rows_zero_s(const Problem &P, const Level &L, int64_t b, int64_t e, const float *R,
                    float *X, hipStream_t st) {
  const Col16 C = col16_of(P, L);
  const unsigned g = grid_for((e - b) * (S / 4));
#define HPCGX_BZ(ST, LI)                                                                      \
  k_bgs32_rows_zero<ST, S, LI><<<g, kBlock, 0, st>>>(b, e, L.n, C, L.zmask, L.vc, L.col,        \
                                                     L.val16, L.diag32, R, X)
  const bool live = P.zero_live & 8;
  if (L.stream && live) HPCGX_BZ(true, true);
  else if (L.stream) HPCGX_BZ(true, false);
  else if (live) HPCGX_BZ(false, true);
  else HPCGX_BZ(false, false);
#undef HPCGX_BZ
  HPCGX_LAUNCHED();
}

// The colour step of a forward sweep that began from X = 0 (only where
// zero_guess_block holds).
void gs_rows_zero(const Problem &P, const Level &L, int s, int64_t b, int64_t e, const float *R,
                  float *X, hipStream_t st) {
  if (s == 4) gs_rows_zero_s<4>(P, L, b, e, R, X, st);
  else if (s == 8) gs_rows_zero_s<8>(P, L, b, e, R, X, st);
  else gs_rows_zero_s<16>(P, L, b, e, R, X, st);
}

void gs_rows(const Problem &P, const Level &L, int s, int64_t b, int64_t e, const float *R,
             float *X, hipStream_t st) {
  if (s == 4) gs_rows_s<4>(P, L, b, e, R, X, st);
  else if (s == 8) gs_rows_s<8>(P, L, b, e, R, X, st);
  else gs_rows_s<16>(P, L, b, e, R, X, st);
}

template <int S>
void restrict_rows_s(const Problem &P, const Level &L, const float *RF, const float *XF,
                     float *RC, hipStream_t st) {
  const Col16 C = col16_of(P, L);
  const bool c16 = L.vc && use_col16(P, L, 0);
  const unsigned g = grid_for(L.nc * (S / 4));
  if (L.stream && c16)
    k_bresid_restrict32_rows<true, S, true><<<g, kBlock, 0, st>>>(
        L.nc, L.ftile, L.rperm, L.n, C, L.vc, L.col, L.val16, L.diag32, RF, XF, RC);
  else if (L.stream)
    k_bresid_restrict32_rows<true, S, false><<<g, kBlock, 0, st>>>(
        L.nc, L.ftile, L.rperm, L.n, C, L.vc, L.col, L.val16, L.diag32, RF, XF, RC);
  else if (c16)
    k_bresid_restrict32_rows<false, S, true><<<g, kBlock, 0, st>>>(
        L.nc, L.ftile, L.rperm, L.n, C, L.vc, L.col, L.val16, L.diag32, RF, XF, RC);
  else
    k_bresid_restrict32_rows<false, S, false><<<g, kBlock, 0, st>>>(
        L.nc, L.ftile, L.rperm, L.n, C, L.vc, L.col, L.val16, L.diag32, RF, XF, RC);
  HPCGX_LAUNCHED();
}

void restrict_rows(const Problem &P, const Level &L, int s, const float *RF, const float *XF,
                   float *RC, hipStream_t st) {
  if (s == 4) restrict_rows_s<4>(P, L, RF, XF, RC, st);
  else if (s == 8) restrict_rows_s<8>(P, L, RF, XF, RC, st);
  else restrict_rows_s<16>(P, L, RF, XF, RC, st);
}

void prolong_rows(const Level &L, int s, const float *XC, float *XF, hipStream_t st) {
  const unsigned g = grid_for(L.nc * (s / 4));
  if (s == 4)
    k_bprolong32_rows<4><<<g, kBlock, 0, st>>>(L.nc, L.ftile, L.rperm, XC, XF);
  else if (s == 8)
    k_bprolong32_rows<8><<<g, kBlock, 0, st>>>(L.nc, L.ftile, L.rperm, XC, XF);
  else
    k_bprolong32_rows<16><<<g, kBlock, 0, st>>>(L.nc, L.ftile, L.rperm, XC, XF);
  HPCGX_LAUNCHED();
}

template <int S>
void spmv_rows_s(const Problem &P, const Level &L, const double *X, double *Y, bool dot,
                 unsigned g, hipStream_t st) {
  const Col16 C = col16_of(P, L);
  const bool h = L.vc && P.half_exact && use_col16(P, L, 0);
  auto &B = P.blk;
  constexpr int W = S / 2;
  if (P.spmv_dot_chunked && B.cpart && B.ncpart >= L.nchunks * W * S) {
    const unsigned G = (unsigned)(L.nchunks * W);
#define HPCGX_BSC(ST, HH)                                                                   \
  k_bspmv_chunks<ST, S, HH><<<G, kBlock, 0, st>>>(L.chunks, L.n, C, L.vc, L.dslot, L.diag,    \
                                                  L.col, L.val, X, Y, dot, B.cpart)
    if (L.stream && h) HPCGX_BSC(true, true);
    else if (L.stream) HPCGX_BSC(true, false);
    else if (h) HPCGX_BSC(false, true);
    else HPCGX_BSC(false, false);
#undef HPCGX_BSC
    HPCGX_LAUNCHED();
    if (dot) {
      const int64_t nparts = L.nchunks * W;
      const unsigned R = (unsigned)std::min<int64_t>(256, nparts);
      k_chunk_bdot<<<R, kBlock, 0, st>>>(nparts, S, B.cpart, B.partials, B.ticket, B.scalars,
                                         kBAlpha);
      HPCGX_LAUNCHED();
    }
    return;
  }
  if (L.stream && h)
    k_bspmv_rows<true, S, true><<<g, kBlock, 0, st>>>(L.n, C, L.vc, L.dslot, L.diag, L.col,
                                                      L.val, X, Y, dot, B.partials, B.ticket,
                                                      B.scalars);
  else if (L.stream)
    k_bspmv_rows<true, S, false><<<g, kBlock, 0, st>>>(L.n, C, L.vc, L.dslot, L.diag, L.col,
                                                       L.val, X, Y, dot, B.partials, B.ticket,
                                                       B.scalars);
  else if (h)
    k_bspmv_rows<false, S, true><<<g, kBlock, 0, st>>>(L.n, C, L.vc, L.dslot, L.diag, L.col,
                                                       L.val, X, Y, dot, B.partials, B.ticket,
                                                       B.scalars);
  else
    k_bspmv_rows<false, S, false><<<g, kBlock, 0, st>>>(L.n, C, L.vc, L.dslot, L.diag, L.col,
                                                        L.val, X, Y, dot, B.partials, B.ticket,
                                                        B.scalars);
  HPCGX_LAUNCHED();
}

void spmv_rows(const Problem &P, const Level &L, int s, const double *X, double *Y, bool dot,
               unsigned g, hipStream_t st) {
  if (s == 4) spmv_rows_s<4>(P, L, X, Y, dot, g, st);
  else if (s == 8) spmv_rows_s<8>(P, L, X, Y, dot, g, st);
  else spmv_rows_s<16>(P, L, X, Y, dot, g, st);
}

// Whether level l's first forward block sweep of a V-cycle can start from an
// unfilled X (every colour step through gs_rows_zero).
bool zero_guess_block(const Problem &P, int l, int ordering, int s) {
  const Level &L = P.L[l];
  if (!P.zero_guess || ordering != HPCGX_ORDER_MULTICOLOR || !rows_kernels(s) || !L.zmask ||
      !L.vc || !L.tbase || !use_col16(P, L, 0) || L.n * s * 4 >= ((int64_t)1 << 32))
    return false;
  for (int c = 0; c < L.ncolors; ++c)
    if (L.color_ptr[c] % kTile) return false;
  return true;
}

// zero_x: X was zero when the sweep began (only with zero_guess_block).
void bsymgs32(Problem &P, int l, int ordering, int s, const float *R, float *X,
              hipStream_t st, bool zero_x = false) {
  const Level &L = P.L[l];
  if (ordering == HPCGX_ORDER_MULTICOLOR && rows_kernels(s)) {
    for (int pass = 0; pass < 2; ++pass)
      for (int q = 0; q < L.ncolors; ++q) {
        const int c = pass == 0 ? q : L.ncolors - 1 - q;
        const int64_t b = L.color_ptr[c], e = L.color_ptr[c + 1];
        if (e <= b) continue;
        if (pass == 0 && zero_x)
          gs_rows_zero(P, L, s, b, e, R, X, st);
        else
          gs_rows(P, L, s, b, e, R, X, st);
      }
  } else if (ordering == HPCGX_ORDER_MULTICOLOR) {
    for (int pass = 0; pass < 2; ++pass)
      for (int q = 0; q < L.ncolors; ++q) {
        const int c = pass == 0 ? q : L.ncolors - 1 - q;
        const int64_t b = L.color_ptr[c], e = L.color_ptr[c + 1];
        if (e <= b) continue;
        const unsigned g = grid_for((e - b) * s);
        if (L.stream)
          k_bgs32<true><<<g, kBlock, 0, st>>>(b, e, L.n, s, L.col, L.val16, L.diag32, R, X);
        else
          k_bgs32<false><<<g, kBlock, 0, st>>>(b, e, L.n, s, L.col, L.val16, L.diag32, R, X);
        HPCGX_LAUNCHED();
      }
  } else {
    for (int pass = 0; pass < 2; ++pass)
      for (int q = 0; q < L.nsched; ++q) {
        const int t = pass == 0 ? q : L.nsched - 1 - q;
        const int64_t cnt = L.level_ptr[t + 1] - L.level_ptr[t];
        k_bgs32_list<<<grid_for(cnt * s), kBlock, 0, st>>>(L.level_rows + L.level_ptr[t], cnt,
                                                             L.n, s, L.col, L.val16, L.diag32,
                                                             R, X);
        HPCGX_LAUNCHED();
      }
  }
}

double read_block_scalars(const Problem &P, int s, std::vector<double> &norm2, hipStream_t st) {
  norm2.resize(s);
  HPCGX_HIP(hipMemcpyAsync(norm2.data(), P.blk.scalars + 4 * s, sizeof(double) * s,
                           hipMemcpyDeviceToHost, st));
  HPCGX_HIP(hipStreamSynchronize(st));
  return 0.0;
}

}  // namespace

void prepare_block(Problem &P, int s) {
  HPCGX_REQUIRE(s >= 1 && s <= 64 && 64 % s == 0, "s must divide 64");
  // one work-item per (row, right-hand side): a launch holds < 2^32 of them
  HPCGX_REQUIRE(P.L[0].n * s < ((int64_t)1 << 32) - kBlock,
                "n * s exceeds the 2^32 work-items of one launch");
  prepare_mixed(P);
  auto &B = P.blk;
  if (B.s >= s) return;
  P.bgraph.reset();  // it holds the old buffers
  B.release(P);
  const int64_t n = P.L[0].n;
  B.R = P.arena.alloc<double>((size_t)n * s);
  B.Pv = P.arena.alloc<double>((size_t)n * s);
  B.AP = P.arena.alloc<double>((size_t)n * s);
  B.R32 = P.arena.alloc<float>((size_t)n * s);
  B.Z32 = P.arena.alloc<float>((size_t)n * s);
  for (int l = 0; l + 1 < P.nlevels; ++l) {
    B.rc32[l] = P.arena.alloc<float>((size_t)P.L[l].nc * s);
    B.xc32[l] = P.arena.alloc<float>((size_t)P.L[l].nc * s);
  }
  B.partials = P.arena.alloc<double>((size_t)kMaxPartials * s);
  if (rows_kernels(s)) {  // k_bspmv_chunks: s/2 workgroups per chunk, s sums each
    B.ncpart = P.L[0].nchunks * (s / 2) * s;
    B.cpart = P.arena.alloc<double>((size_t)B.ncpart);
  }
  B.ticket = P.arena.alloc<unsigned>(1);
  HPCGX_HIP(hipMemset(B.ticket, 0, sizeof(unsigned)));
  B.scalars = P.arena.alloc<double>((size_t)5 * s);
  HPCGX_HIP(hipMemset(B.scalars, 0, sizeof(double) * 5 * s));
  B.s = s;
}

void Problem::BlockState::release(Problem &P) {
  const int64_t n = P.L[0].n;
  if (!s) return;
  P.arena.release(R, sizeof(double) * n * s);
  P.arena.release(Pv, sizeof(double) * n * s);
  P.arena.release(AP, sizeof(double) * n * s);
  P.arena.release(R32, sizeof(float) * n * s);
  P.arena.release(Z32, sizeof(float) * n * s);
  for (int l = 0; l + 1 < P.nlevels; ++l) {
    P.arena.release(rc32[l], sizeof(float) * P.L[l].nc * s);
    P.arena.release(xc32[l], sizeof(float) * P.L[l].nc * s);
  }
  P.arena.release(partials, sizeof(double) * kMaxPartials * s);
  if (cpart) P.arena.release(cpart, sizeof(double) * ncpart);
  cpart = nullptr;
  ncpart = 0;
  P.arena.release(ticket, sizeof(unsigned));
  P.arena.release(scalars, sizeof(double) * 5 * s);
  s = 0;
}

// ComputeMG_ref.cpp:39-63 in FP32 for s vectors (mixed.hip's mg32).
void mg32_block(Problem &P, int l, int ordering, int s, const float *R, float *X,
                hipStream_t st) {
  Level &L = P.L[l];
  // X = 0 (ComputeMG_ref.cpp:42): a fill, or implied by a first forward
  // sweep that reads only the X it has written
  const bool zg = zero_guess_block(P, l, ordering, s);
  if (!zg) {
    k_bfill32<<<grid_for(L.n * s), kBlock, 0, st>>>(X, L.n * s);
    HPCGX_LAUNCHED();
  }
  bsymgs32(P, l, ordering, s, R, X, st, zg);
  if (l + 1 < P.nlevels) {
    float *rc = P.blk.rc32[l], *xc = P.blk.xc32[l];
    const bool rows = rows_kernels(s) && L.rperm;
    const unsigned g = grid_for(L.nc * s);
    if (rows) {
      restrict_rows(P, L, s, R, X, rc, st);
    } else {
      if (L.stream)
        k_bresid_restrict32<true><<<g, kBlock, 0, st>>>(L.nc, L.n, s, L.f2c, L.col, L.val16,
                                                        L.diag32, R, X, rc);
      else
        k_bresid_restrict32<false><<<g, kBlock, 0, st>>>(L.nc, L.n, s, L.f2c, L.col, L.val16,
                                                         L.diag32, R, X, rc);
      HPCGX_LAUNCHED();
    }
    mg32_block(P, l + 1, ordering, s, rc, xc, st);
    if (rows) {
      prolong_rows(L, s, xc, X, st);
    } else {
      k_bprolong32<<<g, kBlock, 0, st>>>(L.nc, s, L.f2c, xc, X);
      HPCGX_LAUNCHED();
    }
    bsymgs32(P, l, ordering, s, R, X, st);
  }
}

void spmv_block(Problem &P, int l, int s, const double *X, double *Y, hipStream_t st) {
  const Level &L = P.L[l];
  if (rows_kernels(s)) {
    spmv_rows(P, L, s, X, Y, false, grid_for(L.n * (s / 2)), st);
    return;
  }
  const unsigned g = grid_for(L.n * s);
  if (L.stream)
    k_bspmv_dot<true><<<g, kBlock, 0, st>>>(L.n, s, L.col, L.val, X, Y, false, nullptr,
                                            nullptr, nullptr);
  else
    k_bspmv_dot<false><<<g, kBlock, 0, st>>>(L.n, s, L.col, L.val, X, Y, false, nullptr,
                                             nullptr, nullptr);
  HPCGX_LAUNCHED();
}

namespace {
// Y = A X for the block CG's fine level; with `dot`, p_k . (A p)_k and alpha.
void block_spmv_dot(Problem &P, int s, const double *x, double *y, bool dot, hipStream_t st) {
  auto &B = P.blk;
  const Level &A = P.L[0];
  const int64_t total = A.n * s;
  if (rows_kernels(s)) {
    // W lanes per row: a persistent grid when it reduces, one work-item per
    // (row, lane) otherwise
    const unsigned grow = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>((A.n * (s / 2) + kBlock - 1) / kBlock, kMaxPartials));
    spmv_rows(P, A, s, x, y, dot, dot ? grow : grid_for(A.n * (s / 2)), st);
    return;
  }
  const unsigned g = dot ? block_reduce_grid(total, s) : grid_for(total);
  if (A.stream)
    k_bspmv_dot<true><<<g, kBlock, 0, st>>>(A.n, s, A.col, A.val, x, y, dot, B.partials,
                                            B.ticket, B.scalars);
  else
    k_bspmv_dot<false><<<g, kBlock, 0, st>>>(A.n, s, A.col, A.val, x, y, dot, B.partials,
                                             B.ticket, B.scalars);
  HPCGX_LAUNCHED();
}

// r = b - A x, r32, |r|^2 (CG.cpp:86-89, and the periodic recompute)
template <int S>
void block_residual_fused_s(Problem &P, const double *Bv, const double *X, hipStream_t st) {
  auto &B = P.blk;
  const Level &L = P.L[0];
  const Col16 C = col16_of(P, L);
  const bool h = L.vc && P.half_exact && use_col16(P, L, 0);
  constexpr int W = S / 2;
  const unsigned G = (unsigned)(L.nchunks * W);
#define HPCGX_BRC(ST, HH)                                                                    \
  k_bspmv_resid_chunks<ST, S, HH><<<G, kBlock, 0, st>>>(L.chunks, L.n, C, L.vc, L.dslot,      \
                                                        L.diag, L.col, L.val, X, Bv, B.R,     \
                                                        B.R32, B.cpart)
  if (L.stream && h) HPCGX_BRC(true, true);
  else if (L.stream) HPCGX_BRC(true, false);
  else if (h) HPCGX_BRC(false, true);
  else HPCGX_BRC(false, false);
#undef HPCGX_BRC
  HPCGX_LAUNCHED();
  const int64_t nparts = L.nchunks * W;
  const unsigned R = (unsigned)std::min<int64_t>(256, nparts);
  k_chunk_bdot<<<R, kBlock, 0, st>>>(nparts, S, B.cpart, B.partials, B.ticket, B.scalars, kBStore);
  HPCGX_LAUNCHED();
}

void block_residual(Problem &P, int s, const double *Bv, const double *X, hipStream_t st) {
  auto &B = P.blk;
  const int64_t total = P.L[0].n * s;
  if (P.resid_fused && rows_kernels(s) && P.spmv_dot_chunked && B.cpart &&
      B.ncpart >= P.L[0].nchunks * (s / 2) * s) {
    if (s == 4) block_residual_fused_s<4>(P, Bv, X, st);
    else if (s == 8) block_residual_fused_s<8>(P, Bv, X, st);
    else block_residual_fused_s<16>(P, Bv, X, st);
    return;
  }
  block_spmv_dot(P, s, X, B.AP, false, st);
  k_bresidual0<<<block_reduce_grid(total, s), kBlock, 0, st>>>(total, s, Bv, B.AP, B.R, B.R32,
                                                               B.partials, B.ticket, B.scalars);
  HPCGX_LAUNCHED();
}

// Whether the block iteration's vector kernels take the kBE-per-thread
// form (s divides 64 and cpart holds their column sums), and whether x's
// step is deferred to the next p update (defer_x, with that form).
bool block_rows_vec(const Problem &P, int s) {
  const auto &B = P.blk;
  const int64_t total = P.L[0].n * s;
  const int64_t nb = (total + kBlock * kBE - 1) / (kBlock * kBE);
  return P.xr_rows && B.cpart && 64 % s == 0 && nb * s <= B.ncpart;
}
bool block_defer_x(const Problem &P, int s) { return P.defer_x && block_rows_vec(P, s); }

// x's deferred step now (before a residual recompute, or after the loop)
void block_step_x(Problem &P, int s, double *X, hipStream_t st) {
  auto &B = P.blk;
  const int64_t total = P.L[0].n * s;
  k_bstep_x<<<grid_for(total), kBlock, 0, st>>>(total, s, X, B.Pv, B.scalars);
  HPCGX_LAUNCHED();
}

// One block iteration without the recompute (CG.cpp:102-124 per column).
void block_iteration(Problem &P, int ordering, int s, double *X, bool first, hipStream_t st) {
  auto &B = P.blk;
  const int64_t total = P.L[0].n * s;
  const unsigned gr = block_reduce_grid(total, s);
  // the kBE-per-thread vector kernels where s divides 64 and cpart holds
  // their column sums
  const int64_t nb = (total + kBlock * kBE - 1) / (kBlock * kBE);
  const bool rows = block_rows_vec(P, s);
  const bool defer = block_defer_x(P, s);
  const unsigned R = (unsigned)std::min<int64_t>(256, nb);
  mg32_block(P, 0, ordering, s, B.R32, B.Z32, st);
  const bool nt = P.block_nt && total >= ((int64_t)1 << 24);
  if (rows) {
    if (nt)
      k_bdot32_rows<true><<<(unsigned)nb, kBlock, 0, st>>>(total, s, B.R, B.Z32, B.cpart);
    else
      k_bdot32_rows<false><<<(unsigned)nb, kBlock, 0, st>>>(total, s, B.R, B.Z32, B.cpart);
    HPCGX_LAUNCHED();
    k_chunk_bdot<<<R, kBlock, 0, st>>>(nb, s, B.cpart, B.partials, B.ticket, B.scalars, kBBeta);
  } else {
    k_bdot32<<<gr, kBlock, 0, st>>>(total, s, B.R, B.Z32, B.partials, B.ticket, B.scalars);
  }
  HPCGX_LAUNCHED();
  if (defer && !first && nt)
    k_bupdate_xp32<true><<<(unsigned)nb, kBlock, 0, st>>>(total, s, X, B.Z32, B.Pv, B.scalars);
  else if (defer && !first)
    k_bupdate_xp32<false><<<(unsigned)nb, kBlock, 0, st>>>(total, s, X, B.Z32, B.Pv, B.scalars);
  else
    k_bupdate_p32<<<grid_for(total), kBlock, 0, st>>>(total, s, B.Z32, B.Pv, B.scalars, first);
  HPCGX_LAUNCHED();
  block_spmv_dot(P, s, B.Pv, B.AP, true, st);
  if (rows && defer) {
    if (nt)
      k_bupdate_xr_rows<false, true><<<(unsigned)nb, kBlock, 0, st>>>(total, s, X, B.R, B.Pv,
                                                                      B.AP, B.R32, B.scalars,
                                                                      B.cpart);
    else
      k_bupdate_xr_rows<false, false><<<(unsigned)nb, kBlock, 0, st>>>(total, s, X, B.R, B.Pv,
                                                                       B.AP, B.R32, B.scalars,
                                                                       B.cpart);
    HPCGX_LAUNCHED();
    k_chunk_bdot<<<R, kBlock, 0, st>>>(nb, s, B.cpart, B.partials, B.ticket, B.scalars, kBStore);
  } else if (rows) {
    if (nt)
      k_bupdate_xr_rows<true, true><<<(unsigned)nb, kBlock, 0, st>>>(total, s, X, B.R, B.Pv,
                                                                     B.AP, B.R32, B.scalars,
                                                                     B.cpart);
    else
      k_bupdate_xr_rows<true, false><<<(unsigned)nb, kBlock, 0, st>>>(total, s, X, B.R, B.Pv,
                                                                      B.AP, B.R32, B.scalars,
                                                                      B.cpart);
    HPCGX_LAUNCHED();
    k_chunk_bdot<<<R, kBlock, 0, st>>>(nb, s, B.cpart, B.partials, B.ticket, B.scalars, kBStore);
  } else {
    k_bupdate_xr<<<gr, kBlock, 0, st>>>(total, s, X, B.R, B.Pv, B.AP, B.R32, B.partials, B.ticket,
                                        B.scalars);
  }
  HPCGX_LAUNCHED();
}

// The captured body of iterations 2.. (rebuilt when its key changes).
hipGraphExec_t block_graph(Problem &P, int ordering, int s, double *X) {
  auto &G = P.bgraph;
  if (G.exec && G.x == X && G.ordering == ordering && P.bgraph_s == s) return G.exec;
  G.reset();
  if (!G.cap) HPCGX_HIP(hipStreamCreateWithFlags(&G.cap, hipStreamNonBlocking));
  HPCGX_HIP(hipStreamBeginCapture(G.cap, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    block_iteration(P, ordering, s, X, false, G.cap);
  } catch (...) {
    (void)hipStreamEndCapture(G.cap, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HPCGX_HIP(hipStreamEndCapture(G.cap, &graph));
  const hipError_t e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HPCGX_HIP(e);
  G.x = X;
  G.ordering = ordering;
  P.bgraph_s = s;
  return G.exec;
}
}  // namespace

// CG.cpp:59-143 for s right-hand sides with the FP32 V-cycle (cg_mixed's
// arithmetic for each of them).  Iterates until every system reaches the
// tolerance or max_iter; normr/normr0 per right-hand side.  Iterations 2..
// replay one captured graph (unless the "mixed_graph" knob is off).
void cg_block_mixed(Problem &P, int ordering, int s, const double *Bv, double *X, int max_iter,
                    double tol, int recompute, int *niters, double *normr, double *normr0,
                    double *times, hipStream_t st) {
  prepare_block(P, s);
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<double> n2;
  *niters = 0;
  block_residual(P, s, Bv, X, st);  // :86-89 (p = x, Ap = A p, r = b - Ap)
  read_block_scalars(P, s, n2, st);
  for (int k = 0; k < s; ++k) normr0[k] = normr[k] = std::sqrt(n2[k]);
  auto done = [&]() {
    for (int k = 0; k < s; ++k)
      if (normr[k] / normr0[k] > tol) return false;
    return true;
  };
  const bool check_each = tol > 0.0;
  for (int it = 1; it <= max_iter && !done(); ++it) {
    if (it >= 2 && P.use_mixed_graph && !P.klog.on)
      HPCGX_HIP(hipGraphLaunch(block_graph(P, ordering, s, X), st));
    else
      block_iteration(P, ordering, s, X, it == 1, st);
    if (recompute > 0 && it % recompute == 0) {
      if (block_defer_x(P, s)) {
        // the pending x step now; alpha then reads 0 in the next p update,
        // whose x step (x + 0 p) leaves x as it is
        block_step_x(P, s, X, st);
        launch_fill(P.blk.scalars + 2 * s, s, 0.0, st);
      }
      block_residual(P, s, Bv, X, st);
    }
    *niters = it;
    if (check_each || it == max_iter) {
      read_block_scalars(P, s, n2, st);
      for (int k = 0; k < s; ++k) normr[k] = std::sqrt(n2[k]);
    }
  }
  // the last iteration's deferred x step (none after a recompute, which took
  // it, or before the first iteration)
  if (block_defer_x(P, s) && *niters > 0 && !(recompute > 0 && *niters % recompute == 0))
    block_step_x(P, s, X, st);
  if (!check_each) {
    read_block_scalars(P, s, n2, st);
    for (int k = 0; k < s; ++k) normr[k] = std::sqrt(n2[k]);
  }
  HPCGX_HIP(hipStreamSynchronize(st));
  if (times)
    times[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int k = 0; k < s; ++k)
    if (!std::isfinite(normr[k])) throw Error(HPCGX_EINVAL, "CG residual is not finite");
}

}  // namespace hpcgx
