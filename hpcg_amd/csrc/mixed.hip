// mixed.hip -- the mixed-precision preconditioner: the V-cycle in FP32 with
// matrix values stored in FP16, under the FP64 outer PCG.
//
// Storage per level (built on first use): val16 = the ELL values rounded to
// FP16 (double -> float -> half, round to nearest even each time) with the
// diagonal slot set to 0, and diag32 = the diagonal in FP32, so that a
// diagonal scaled beyond the FP16 range (TestCG multiplies it by up to
// 1.1e7) stays exact to FP32.  Column ids are shared with the FP64 ELL, so
// the V-cycle streams 6 bytes per stored entry instead of 12.
//
// Row arithmetic (no fused multiply-add, correctly rounded division):
//   Gauss-Seidel  sum = r_i; sum -= h_ij * x_j over the slots (the diagonal
//                 and unused slots hold 0); x_i = sum / d_i
//   A*x           s = sum_j h_ij * x_j over the slots; s += d_i * x_i
// The outer PCG (CG.cpp:59-143 of the reference) stays in FP64: dot
// products, alpha, beta, x and r, and A p with FP64 products and sums (its
// off-diagonal values read from the FP16 copy, the diagonal from the FP64
// array: the FP64 SpMV while those values are exact in FP16, as -1 is);
// every `recompute`-th iteration r = b - A x is recomputed with the FP64
// matrix instead of updated.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "internal.hpp"

namespace hpcgx {
namespace {

typedef _Float16 half_t;

inline unsigned grid_for(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}
__device__ inline int64_t gtid() {
  return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}

template <bool STREAM, class T>
__device__ inline T mload(const T *p) {
  if (STREAM) return __builtin_nontemporal_load(p);
  return *p;
}

// FP16 copy of the off-diagonal values (diagonal slot zeroed) and the FP32
// diagonal.  flags |= 1: an off-diagonal value overflows FP16; flags |= 2: an
// off-diagonal value is not exact in FP16 (the FP16-read outer SpMV then
// falls back to the FP64 values).
__global__ __launch_bounds__(kBlock) void k_make_half(
    int64_t n, const double *__restrict__ val, const uint8_t *__restrict__ dslot,
    half_t *__restrict__ val16, float *__restrict__ diag32, int *flags) {
  const int64_t i = gtid();
  if (i >= n) return;
  const int ds = dslot[i];
  bool bad = false, inexact = false;
  for (int j = 0; j < kSlots; ++j) {
    const int64_t k = (int64_t)j * n + i;
    const double v = val[k];
    const float f = (float)v;
    const half_t h = (j == ds) ? (half_t)0.0f : (half_t)f;
    bad |= (j != ds) && __builtin_isinf((float)h) && !__builtin_isinf(f);
    inexact |= (j != ds) && (double)(float)h != v;
    val16[k] = h;
  }
  diag32[i] = (float)val[(int64_t)ds * n + i];
  if (bad || inexact) atomicOr(flags, (bad ? 1 : 0) | (inexact ? 2 : 0));
}

__global__ __launch_bounds__(kBlock) void k_diag32(int64_t n, const double *d,
                                                    float *diag32) {
  const int64_t i = gtid();
  if (i < n) diag32[i] = (float)d[i];
}

__global__ __launch_bounds__(kBlock) void k_fill32(float *x, int64_t n) {
  const int64_t i = gtid();
  if (i < n) x[i] = 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double *in,
                                                    float *out) {
  const int64_t i = gtid();
  if (i < n) out[i] = (float)in[i];
}

__global__ __launch_bounds__(kBlock) void k_to_f64(int64_t n, const float *in,
                                                    double *out) {
  const int64_t i = gtid();
  if (i < n) out[i] = (double)in[i];
}

template <bool STREAM>
__device__ inline float gs32_row(int64_t i, int64_t n,
                                const int32_t *__restrict__ col,
                                const half_t *__restrict__ val16,
                                const float *__restrict__ diag32,
                                const float *__restrict__ r, float *x) {
  float sum = r[i];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t k = (int64_t)j * n + i;
    sum -= (float)mload<STREAM>(val16 + k) * x[mload<STREAM>(col + k)];
  }
  const float v = sum / diag32[i];
  x[i] = v;
  return v;
}

template <bool STREAM>
__device__ inline float row_dot32(int64_t i, int64_t n,
                                  const int32_t *__restrict__ col,
                                  const half_t *__restrict__ val16,
                                  const float *__restrict__ diag32,
                                  const float *__restrict__ x) {
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t k = (int64_t)j * n + i;
    s += (float)mload<STREAM>(val16 + k) * x[mload<STREAM>(col + k)];
  }
  s += diag32[i] * x[i];
  return s;
}

// 16-bit column ids (Level::col16): the caller's wavefront lies in one tile.
__device__ inline int tile_of(int64_t i) {
  return __builtin_amdgcn_readfirstlane((int)(i / kTile));
}

// Entry of a narrow tile packed in one 32-bit word (Level::vc):
// high half = 16-bit column offset, low half = the FP16 value.  One load per
// slot instead of two: the FP32 colour step is bound by the number of lane
// addresses the memory pipeline processes, not by bytes.
__device__ inline half_t vc_val(uint32_t w) {
  return __builtin_bit_cast(half_t, (uint16_t)(w & 0xffffu));
}

template <bool STREAM>
__device__ inline float gs32_row16(const Col16 &C, const uint32_t *__restrict__ vc,
                                  int64_t i, int64_t n, const int32_t *__restrict__ col,
                                  const half_t *__restrict__ val16,
                                  const float *__restrict__ diag32,
                                  const float *__restrict__ r, float *x) {
  const int t = tile_of(i);
  const int info = C.info[t];
  if (info == kTile32) return gs32_row<STREAM>(i, n, col, val16, diag32, r, x);
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
  float sum = r[i];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const uint32_t w = mload<STREAM>(vc + (int64_t)j * n + i);
    const int32_t c = base[j] + (int32_t)(w >> 16);
    sum -= (float)vc_val(w) * x[c];
  }
  const float v = sum / diag32[i];
  x[i] = v;
  return v;
}

template <bool STREAM>
__device__ inline float row_dot32_16(const Col16 &C, const uint32_t *__restrict__ vc,
                                     int64_t i, int64_t n, const int32_t *__restrict__ col,
                                     const half_t *__restrict__ val16,
                                     const float *__restrict__ diag32,
                                     const float *__restrict__ x) {
  const int t = tile_of(i);
  const int info = C.info[t];
  if (info == kTile32) return row_dot32<STREAM>(i, n, col, val16, diag32, x);
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const uint32_t w = mload<STREAM>(vc + (int64_t)j * n + i);
    const int32_t c = base[j] + (int32_t)(w >> 16);
    s += (float)vc_val(w) * x[c];
  }
  s += diag32[i] * x[i];
  return s;
}

// Whether row i's wavefront can share x windows between slots 3m and 3m + 2
// (kernels.hip's pairs_tile on the tile-major bases): all offset words of the
// tile affine, slot 3m + 2's base one above slot 3m's for every m, and all
// 64 rows below `end`.  Uniform per wavefront.
__device__ inline bool pairs_tile32(const Col16 &C, int t, unsigned am, int64_t end) {
  if (am != (1u << kSlotWords) - 1u || (int64_t)(t + 1) * kTile > end) return false;
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
  bool ok = true;
#pragma unroll
  for (int m = 0; m < kSlots / 3; ++m) ok = ok && base[3 * m + 2] == base[3 * m] + 1;
  return ok;
}

// FP32 colour step of an 8-bit tile (Level::off8, Level::v16p): per row
// kSlotWords offset words and kPairWords value words, 3 bytes per slot, in
// gs32_row16's order of terms.  Other tiles take gs32_row16.  PAIRS: slot
// 3m + 2's x values from slot 3m's where pairs_tile32 holds.
template <bool STREAM, bool PAIRS = false>
__device__ inline float gs32_row8(const Col16 &C, const uint32_t *__restrict__ vp,
                                 const uint32_t *__restrict__ vc, int64_t i, int64_t n,
                                 const int32_t *__restrict__ col,
                                 const half_t *__restrict__ val16,
                                 const float *__restrict__ diag32,
                                 const float *__restrict__ r, float *x, int64_t end = 0) {
  const int t = tile_of(i);
  const int info = C.info[t];
  if (info != kTile8 || !C.use8) return gs32_row16<STREAM>(C, vc, i, n, col, val16, diag32, r, x);
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
  const unsigned am = C.amask ? C.amask[t] : 0u;
  uint32_t o8[kSlotWords], pv[kPairWords];
#pragma unroll
  for (int g = 0; g < kSlotWords; ++g) o8[g] = off8_word<STREAM>(C.off8, am, g, n, i);
#pragma unroll
  for (int w = 0; w < kPairWords; ++w) pv[w] = mload<STREAM>(vp + (int64_t)w * n + i);
  float sum = r[i];
  if (PAIRS && pairs_tile32(C, t, am, end)) {
    // kernels.hip's affine_pairs_row in FP32: slot 3m + 2's x from slot 3m's,
    // one lane down; lane 63's from one address for the wavefront
    const int32_t pos = (int32_t)(i % kTile);
#pragma unroll
    for (int m = 0; m < kSlots / 3; ++m) {
      const int32_t b0 = base[3 * m];
      const float x0 = x[b0 + pos];
      const float x1 = x[base[3 * m + 1] + pos];
      const float up = __shfl_down(x0, 1, 64);
      const float last = x[b0 + kTile];
      const float x2 = pos == kTile - 1 ? last : up;
      const float xs[3] = {x0, x1, x2};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int j = 3 * m + q;
        const half_t a = __builtin_bit_cast(half_t, (uint16_t)(pv[j / 2] >> (16 * (j % 2))));
        sum -= (float)a * xs[q];
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int32_t c = base[j] + (int32_t)((o8[j / 4] >> (8 * (j % 4))) & 0xffu);
      const half_t a = __builtin_bit_cast(half_t, (uint16_t)(pv[j / 2] >> (16 * (j % 2))));
      sum -= (float)a * x[c];
    }
  }
  const float v = sum / diag32[i];
  x[i] = v;
  return v;
}

// row_dot32_16 on an 8-bit tile from the FP16 value pairs (gs32_row8's
// reads); other tiles take row_dot32_16.
template <bool STREAM>
__device__ inline float row_dot32_8(const Col16 &C, const uint32_t *__restrict__ vp,
                                    const uint32_t *__restrict__ vc, int64_t i, int64_t n,
                                    const int32_t *__restrict__ col,
                                    const half_t *__restrict__ val16,
                                    const float *__restrict__ diag32,
                                    const float *__restrict__ x) {
  const int t = tile_of(i);
  const int info = C.info[t];
  if (info != kTile8 || !C.use8) return row_dot32_16<STREAM>(C, vc, i, n, col, val16, diag32, x);
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
  const unsigned am = C.amask ? C.amask[t] : 0u;
  uint32_t o8[kSlotWords], pv[kPairWords];
#pragma unroll
  for (int g = 0; g < kSlotWords; ++g) o8[g] = off8_word<STREAM>(C.off8, am, g, n, i);
#pragma unroll
  for (int w = 0; w < kPairWords; ++w) pv[w] = mload<STREAM>(vp + (int64_t)w * n + i);
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int32_t c = base[j] + (int32_t)((o8[j / 4] >> (8 * (j % 4))) & 0xffu);
    const half_t a = __builtin_bit_cast(half_t, (uint16_t)(pv[j / 2] >> (16 * (j % 2))));
    s += (float)a * x[c];
  }
  s += diag32[i] * x[i];
  return s;
}

// The first forward sweep of an FP32 V-cycle level from x = 0: the FP32
// form of kernels.hip's zero-x colour step (see gs_row_zero there).  Slots
// that Level::zmask leaves out of the tile, and x at or beyond `begin`,
// are read through buffer views whose range check returns 0 without a
// memory access, so the row's loads still go out together.  Each row keeps
// gs32_row8's (8-bit tiles with FP16 value pairs), gs32_row16's (packed
// words) or gs32_row's (32-bit ids) term order; skipped terms are a * 0.
typedef __amdgpu_buffer_rsrc_t BufView;
constexpr int kBufWord3 = 0x00020000;  // buffer descriptor word 3 for gfx9 raw loads

__device__ inline BufView buf_view(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kBufWord3);
}
template <bool STREAM>
__device__ inline uint32_t buf_u32(BufView v, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(v, off, 0, STREAM ? 2 : 0);
}
__device__ inline float buf_f32(BufView v, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(v, off, 0, 0));
}
template <bool STREAM>
__device__ inline uint32_t buf_u16(BufView v, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b16(v, off, 0, STREAM ? 2 : 0);
}

template <bool STREAM>
__device__ inline void gs32_row_zero(const Col16 &C, uint32_t zm, int64_t begin,
                                     const uint32_t *__restrict__ vp,
                                     const uint32_t *__restrict__ vc, int64_t i, int64_t n,
                                     const int32_t *__restrict__ col,
                                     const half_t *__restrict__ val16,
                                     const float *__restrict__ diag32,
                                     const float *__restrict__ r, float *x) {
  const int t = tile_of(i);
  const int info = C.info[t];
  const int64_t tb = (int64_t)t * kTile;
  const uint32_t lane = (uint32_t)(i - tb);
  const uint32_t xlim = (uint32_t)begin * 4u;
  float sum = r[i];
  if (info == kTile8 && C.use8 && vp) {
    const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
    const unsigned am = C.amask ? C.amask[t] : 0u;
    uint32_t o8[kSlotWords], pv[kPairWords];
#pragma unroll
    for (int g = 0; g < kSlotWords; ++g) {
      const bool any = (zm >> (4 * g)) & 0xfu;
      o8[g] = ((am >> g) & 1u)
                  ? lane * 0x01010101u
                  : buf_u32<STREAM>(buf_view(C.off8 + (int64_t)g * n + tb, any ? 256u : 0u),
                                    lane * 4u);
    }
#pragma unroll
    for (int w = 0; w < kPairWords; ++w) {
      const bool any = (zm >> (2 * w)) & 3u;
      pv[w] = buf_u32<STREAM>(buf_view(vp + (int64_t)w * n + tb, any ? 256u : 0u), lane * 4u);
    }
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const bool live = (zm >> j) & 1u;
      const int32_t c = base[j] + (int32_t)((o8[j / 4] >> (8 * (j % 4))) & 0xffu);
      const half_t a = __builtin_bit_cast(half_t, (uint16_t)(pv[j / 2] >> (16 * (j % 2))));
      sum -= (float)a * buf_f32(buf_view(x, live ? xlim : 0u), (uint32_t)c * 4u);
    }
  } else if (info != kTile32) {
    const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const bool live = (zm >> j) & 1u;
      const uint32_t w =
          buf_u32<STREAM>(buf_view(vc + (int64_t)j * n + tb, live ? 256u : 0u), lane * 4u);
      const int32_t c = base[j] + (int32_t)(w >> 16);
      sum -= (float)vc_val(w) * buf_f32(buf_view(x, live ? xlim : 0u), (uint32_t)c * 4u);
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const bool live = (zm >> j) & 1u;
      const int32_t c = (int32_t)buf_u32<STREAM>(
          buf_view(col + (int64_t)j * n + tb, live ? 256u : 0u), lane * 4u);
      const half_t a = __builtin_bit_cast(
          half_t, (uint16_t)buf_u16<STREAM>(buf_view(val16 + (int64_t)j * n + tb,
                                                     live ? 128u : 0u),
                                            lane * 2u));
      sum -= (float)a * buf_f32(buf_view(x, live ? xlim : 0u), (uint32_t)c * 4u);
    }
  }
  x[i] = sum / diag32[i];
}

// gs32_row_zero over the live slots only (tune key "zero_live"): Level::
// zmask's set bits in ascending slot order, CH per chunk, with the
// chunk's unused entries read through empty views (0 * 0).  Each live slot
// loads its own offset word and value word, so the loads of dead slots are
// not issued.  CLS: 0 32-bit ids, 1 packed words, 2 8-bit offsets with FP16
// value pairs, 3 the same with every offset word affine (no offset loads);
// the tile class is decided once outside the loop, and the chunk is
// straight code (an affine word is a select, not a branch), so its loads
// go out together.
template <bool STREAM, int CLS, int CH>
__device__ inline float zero_live_sum32(const Col16 &C, uint32_t zm, int t, int64_t tb,
                                        uint32_t lane, uint32_t xlim, int64_t n,
                                        const uint32_t *__restrict__ vp,
                                        const uint32_t *__restrict__ vc,
                                        const int32_t *__restrict__ col,
                                        const half_t *__restrict__ val16, const float *x,
                                        float sum) {
  const int32_t *__restrict__ base = C.tbase + (int64_t)t * kBaseStride;
  const unsigned am = (CLS == 2 && C.amask) ? C.amask[t] : 0u;  // affine offset words
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
    // the chunk's index and value words, then its x values, then the sums
    uint32_t wa[CH], wb[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int j = js[q];
      if (CLS >= 2) {
        // an affine offset word (every one for CLS 3) is not loaded
        const bool aff = CLS == 3 || ((am >> (j >> 2)) & 1u);
        wa[q] = CLS == 3 ? 0u
                         : buf_u32<STREAM>(buf_view(C.off8 + (int64_t)(j >> 2) * n + tb,
                                                    ok[q] && !aff ? 256u : 0u),
                                           lane * 4u);
        wb[q] = buf_u32<STREAM>(buf_view(vp + (int64_t)(j >> 1) * n + tb, ok[q] ? 256u : 0u),
                                lane * 4u);
      } else if (CLS == 1) {
        wa[q] = buf_u32<STREAM>(buf_view(vc + (int64_t)j * n + tb, ok[q] ? 256u : 0u), lane * 4u);
      } else {
        wa[q] = buf_u32<STREAM>(buf_view(col + (int64_t)j * n + tb, ok[q] ? 256u : 0u), lane * 4u);
        wb[q] = buf_u16<STREAM>(buf_view(val16 + (int64_t)j * n + tb, ok[q] ? 128u : 0u),
                                lane * 2u);
      }
    }
    float a[CH], xc[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int j = js[q];
      int32_t c;
      if (CLS >= 2) {
        const bool aff = CLS == 3 || ((am >> (j >> 2)) & 1u);
        const uint32_t o = aff ? lane * 0x01010101u : wa[q];
        c = base[j] + (int32_t)((o >> (8 * (j & 3))) & 0xffu);
        a[q] = (float)__builtin_bit_cast(half_t, (uint16_t)(wb[q] >> (16 * (j & 1))));
      } else if (CLS == 1) {
        c = base[j] + (int32_t)(wa[q] >> 16);
        a[q] = (float)vc_val(wa[q]);
      } else {
        c = (int32_t)wa[q];
        a[q] = (float)__builtin_bit_cast(half_t, (uint16_t)wb[q]);
      }
      xc[q] = buf_f32(buf_view(x, ok[q] ? xlim : 0u), (uint32_t)c * 4u);
    }
#pragma unroll
    for (int q = 0; q < CH; ++q) sum -= a[q] * xc[q];
  }
  return sum;
}

template <bool STREAM, int CH>
__device__ inline void gs32_row_zero_live(const Col16 &C, uint32_t zm, int64_t begin,
                                          const uint32_t *__restrict__ vp,
                                          const uint32_t *__restrict__ vc, int64_t i, int64_t n,
                                          const int32_t *__restrict__ col,
                                          const half_t *__restrict__ val16,
                                          const float *__restrict__ diag32,
                                          const float *__restrict__ r, float *x) {
  // every wavefront's rows lie in one tile (colour blocks start on tiles in
  // this layout, and workgroups are whole wavefronts), so the tile and its
  // mask are uniform: the slot loop runs on scalar registers
  const int t = __builtin_amdgcn_readfirstlane(tile_of(i));
  zm = __builtin_amdgcn_readfirstlane(zm);
  const int info = C.info[t];
  const int64_t tb = (int64_t)t * kTile;
  const uint32_t lane = (uint32_t)(i - tb);
  const uint32_t xlim = (uint32_t)begin * 4u;
  float sum = r[i];
  if (info == kTile8 && C.use8 && vp && C.amask && C.amask[t] == (1u << kSlotWords) - 1u)
    sum = zero_live_sum32<STREAM, 3, CH>(C, zm, t, tb, lane, xlim, n, vp, vc, col, val16, x, sum);
  else if (info == kTile8 && C.use8 && vp)
    sum = zero_live_sum32<STREAM, 2, CH>(C, zm, t, tb, lane, xlim, n, vp, vc, col, val16, x, sum);
  else if (info != kTile32)
    sum = zero_live_sum32<STREAM, 1, CH>(C, zm, t, tb, lane, xlim, n, vp, vc, col, val16, x, sum);
  else
    sum = zero_live_sum32<STREAM, 0, CH>(C, zm, t, tb, lane, xlim, n, vp, vc, col, val16, x, sum);
  x[i] = sum / diag32[i];
}

// v16p[w*n + i] = bits of val16[(2w)*n + i] | bits of val16[(2w+1)*n + i] << 16
// (0 above the last slot), one thread per row.
__global__ __launch_bounds__(kBlock) void k_pack_v16p(int64_t n, const half_t *__restrict__ val16,
                                                      uint32_t *__restrict__ vp) {
  const int64_t i = gtid();
  if (i >= n) return;
#pragma unroll
  for (int w = 0; w < kPairWords; ++w) {
    const uint32_t lo = __builtin_bit_cast(uint16_t, val16[(int64_t)(2 * w) * n + i]);
    const uint32_t hi =
        2 * w + 1 < kSlots ? __builtin_bit_cast(uint16_t, val16[(int64_t)(2 * w + 1) * n + i]) : 0u;
    vp[(int64_t)w * n + i] = lo | (hi << 16);
  }
}

// The hardware hands workgroup b to XCD b % 8; with `xcd` XCD k takes the
// k-th eighth of the rows in order (its L2 then holds its own x window).
__device__ inline int64_t xcd_block(int64_t b, int64_t G, bool on) {
  if (!on) return b;
  const int64_t G8 = G - G % 8;
  if (b >= G8) return b;
  return (b % 8) * (G8 / 8) + b / 8;
}

template <bool STREAM, bool FINEST, bool PAIRS = false>
__global__ __launch_bounds__(kBlock) void k_gs32_range8(
    int64_t begin, int64_t end, int64_t n, Col16 C, const uint32_t *__restrict__ vp,
    const uint32_t *__restrict__ vc, const int32_t *__restrict__ col,
    const half_t *__restrict__ val16, const float *__restrict__ diag32,
    const float *__restrict__ r, float *x, bool xcd) {
  const int64_t i = begin + xcd_block(blockIdx.x, gridDim.x, xcd) * kBlock + threadIdx.x;
  if (i < end) gs32_row8<STREAM, PAIRS>(C, vp, vc, i, n, col, val16, diag32, r, x, end);
}

// LIVE: the live slots only (gs32_row_zero_live), else every slot (gs32_row_zero)
template <bool STREAM, bool FINEST, bool LIVE = false>
__global__ __launch_bounds__(kBlock) void k_gs32_zero(
    int64_t begin, int64_t end, int64_t n, Col16 C, const uint32_t *__restrict__ zmask,
    const uint32_t *__restrict__ vp, const uint32_t *__restrict__ vc,
    const int32_t *__restrict__ col, const half_t *__restrict__ val16,
    const float *__restrict__ diag32, const float *__restrict__ r, float *x, bool xcd) {
  const int64_t i = begin + xcd_block(blockIdx.x, gridDim.x, xcd) * kBlock + threadIdx.x;
  if (i >= end) return;
  if (LIVE)
    gs32_row_zero_live<STREAM, 8>(C, zmask[tile_of(i)], begin, vp, vc, i, n, col, val16, diag32,
                                  r, x);
  else
    gs32_row_zero<STREAM>(C, zmask[tile_of(i)], begin, vp, vc, i, n, col, val16, diag32, r, x);
}

// k_gs32_range8 for the V-cycle's last sweep (the finest level's backward
// post-smoother): wavefront k of the launch (rows begin + 64 k ..) also
// stores the sum of r64_i * (double)(new x_i) over its rows to dpart[k], so
// the mixed CG's r.z (CG.cpp:110, r in FP64, z widened) needs no pass over
// r and z of its own.  One partial per wavefront, a 64-lane shuffle tree.
template <bool STREAM, bool PAIRS = false>
__global__ __launch_bounds__(kBlock) void k_gs32_rz8(
    int64_t begin, int64_t end, int64_t n, Col16 C, const uint32_t *__restrict__ vp,
    const uint32_t *__restrict__ vc, const int32_t *__restrict__ col,
    const half_t *__restrict__ val16, const float *__restrict__ diag32,
    const float *__restrict__ r, float *x, bool xcd, const double *__restrict__ r64,
    double *__restrict__ dpart) {
  const int64_t t = xcd_block(blockIdx.x, gridDim.x, xcd) * kBlock + threadIdx.x;
  const int64_t i = begin + t;
  double sp = 0.0;
  if (i < end) {
    const float z = gs32_row8<STREAM, PAIRS>(C, vp, vc, i, n, col, val16, diag32, r, x, end);
    sp = mload<STREAM>(r64 + i) * (double)z;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sp += __shfl_down(sp, off, 64);
  if ((threadIdx.x & 63) == 0 && i < end) dpart[t >> 6] = sp;
}

// vc[k] = col16[k] << 16 | bits of val16[k], one thread per row (a launch
// holds fewer than 2^32 work-items; 27 n can exceed that).
__global__ __launch_bounds__(kBlock) void k_pack_vc(int64_t n, const uint16_t *__restrict__ off,
                                                    const half_t *__restrict__ val16,
                                                    uint32_t *__restrict__ vc) {
  const int64_t i = gtid();
  if (i >= n) return;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t k = (int64_t)j * n + i;
    vc[k] = ((uint32_t)off[k] << 16) | __builtin_bit_cast(uint16_t, val16[k]);
  }
}

template <bool STREAM, bool FINEST>
__global__ __launch_bounds__(kBlock) void k_gs32_range16(
    int64_t begin, int64_t end, int64_t n, Col16 C, const uint32_t *__restrict__ vc,
    const int32_t *__restrict__ col, const half_t *__restrict__ val16,
    const float *__restrict__ diag32, const float *__restrict__ r, float *x) {
  const int64_t i = begin + gtid();
  if (i < end) gs32_row16<STREAM>(C, vc, i, n, col, val16, diag32, r, x);
}

template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_spmv32_16(
    const int64_t *__restrict__ chunks, int64_t n, Col16 C, const uint32_t *__restrict__ vc,
    const int32_t *__restrict__ col, const half_t *__restrict__ val16,
    const float *__restrict__ diag32, const float *__restrict__ x,
    float *__restrict__ y) {
  const int64_t b = blockIdx.x;
  const int64_t i = chunks[2 * b] + threadIdx.x;
  if (i < chunks[2 * b + 1])
    y[i] = row_dot32_16<STREAM>(C, vc, i, n, col, val16, diag32, x);
}

template <bool STREAM, bool FINEST>
__global__ __launch_bounds__(kBlock) void k_gs32_range(
    int64_t begin, int64_t end, int64_t n, const int32_t *__restrict__ col,
    const half_t *__restrict__ val16, const float *__restrict__ diag32,
    const float *__restrict__ r, float *x) {
  const int64_t i = begin + gtid();
  if (i < end) gs32_row<STREAM>(i, n, col, val16, diag32, r, x);
}

__global__ __launch_bounds__(kBlock) void k_gs32_list(
    const int32_t *__restrict__ rows, int64_t count, int64_t n,
    const int32_t *__restrict__ col, const half_t *__restrict__ val16,
    const float *__restrict__ diag32, const float *__restrict__ r, float *x) {
  const int64_t k = gtid();
  if (k < count) gs32_row<false>(rows[k], n, col, val16, diag32, r, x);
}

template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_spmv32(
    const int64_t *__restrict__ chunks, int64_t n,
    const int32_t *__restrict__ col, const half_t *__restrict__ val16,
    const float *__restrict__ diag32, const float *__restrict__ x,
    float *__restrict__ y) {
  const int64_t b = blockIdx.x;
  const int64_t i = chunks[2 * b] + threadIdx.x;
  if (i < chunks[2 * b + 1]) y[i] = row_dot32<STREAM>(i, n, col, val16, diag32, x);
}

template <bool STREAM>
__global__ __launch_bounds__(kBlock) void k_resid_restrict32(
    int64_t nc, int64_t n, const int32_t *__restrict__ f2c,
    const int32_t *__restrict__ col, const half_t *__restrict__ val16,
    const float *__restrict__ diag32, const float *__restrict__ rf,
    const float *__restrict__ xf, float *__restrict__ rc) {
  const int64_t c = gtid();
  if (c < nc) {
    const int64_t f = f2c[c];
    rc[c] = rf[f] - row_dot32<STREAM>(f, n, col, val16, diag32, xf);
  }
}

__global__ __launch_bounds__(kBlock) void k_prolong32(
    int64_t nc, const int32_t *__restrict__ f2c, const float *__restrict__ xc,
    float *__restrict__ xf) {
  const int64_t c = gtid();
  if (c < nc) xf[f2c[c]] += xc[c];
}

// The transfers in fine order (Level::ftile, rperm): thread t owns fine row
// fine_row(ftile, t) (whole tiles per wavefront, packed words when C16) and
// coarse row rperm[t].
// C16 = 2: FP16 value pairs on 8-bit tiles (row_dot32_8).
template <bool STREAM, int C16>
__global__ __launch_bounds__(kBlock) void k_resid_restrict32_f(
    int64_t nc, const int32_t *__restrict__ ftile, const int32_t *__restrict__ rperm, int64_t n,
    Col16 C, const uint32_t *__restrict__ vc, const uint32_t *__restrict__ vp,
    const int32_t *__restrict__ col, const half_t *__restrict__ val16,
    const float *__restrict__ diag32, const float *__restrict__ rf,
    const float *__restrict__ xf, float *__restrict__ rc) {
  const int64_t t = gtid();
  if (t >= nc) return;
  const int64_t f = fine_row(ftile, t);
  const float ax = C16 == 2   ? row_dot32_8<STREAM>(C, vp, vc, f, n, col, val16, diag32, xf)
                   : C16 == 1 ? row_dot32_16<STREAM>(C, vc, f, n, col, val16, diag32, xf)
                              : row_dot32<STREAM>(f, n, col, val16, diag32, xf);
  rc[rperm[t]] = rf[f] - ax;
}

__global__ __launch_bounds__(kBlock) void k_prolong32_f(int64_t nc,
                                                        const int32_t *__restrict__ ftile,
                                                        const int32_t *__restrict__ rperm,
                                                        const float *__restrict__ xc,
                                                        float *__restrict__ xf) {
  const int64_t t = gtid();
  if (t < nc) xf[fine_row(ftile, t)] += xc[rperm[t]];
}

// The FP16-pair colour step applies (8-bit offsets and FP16 value pairs).
bool pairs_step(const Problem &P, const Level &L, int64_t begin) {
  return use_col16(P, L, begin) && L.v16p && P.use_pairs && P.use_c8 && L.byte_tiles > 0;
}

// With dpart (only where pairs_step holds), also the r.z partials of
// k_gs32_rz8 from the FP64 residual r64.
void gs32_step(const Problem &P, const Level &L, int64_t begin, int64_t end,
               const float *r, float *x, hipStream_t s, double *dpart = nullptr,
               const double *r64 = nullptr) {
  if (end <= begin) return;
  const unsigned g = grid_for(end - begin);
  if (dpart && !pairs_step(P, L, begin))
    throw Error(HPCGX_EINVAL, "FP32 colour step with r.z partials needs the FP16-pair layout");
  if (pairs_step(P, L, begin)) {
    const Col16 C = col16_of(P, L);
    const bool xp = L.finest && (P.x_pairs & 4);  // the x-pair steps (finest level)
    if (dpart && xp) {
      k_gs32_rz8<true, true><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.v16p, L.vc, L.col,
                                                  L.val16, L.diag32, r, x, L.xcd, r64, dpart);
    } else if (xp) {
      k_gs32_range8<true, true, true><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.v16p, L.vc,
                                                           L.col, L.val16, L.diag32, r, x, L.xcd);
    } else if (dpart) {
      if (L.finest || L.stream)
        k_gs32_rz8<true><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.v16p, L.vc, L.col, L.val16,
                                              L.diag32, r, x, L.xcd, r64, dpart);
      else
        k_gs32_rz8<false><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.v16p, L.vc, L.col,
                                               L.val16, L.diag32, r, x, L.xcd, r64, dpart);
    } else if (L.finest)
      k_gs32_range8<true, true><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.v16p, L.vc, L.col,
                                                     L.val16, L.diag32, r, x, L.xcd);
    else if (L.stream)
      k_gs32_range8<true, false><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.v16p, L.vc, L.col,
                                                      L.val16, L.diag32, r, x, L.xcd);
    else
      k_gs32_range8<false, false><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.v16p, L.vc, L.col,
                                                       L.val16, L.diag32, r, x, L.xcd);
    HPCGX_LAUNCHED();
    return;
  }
  if (use_col16(P, L, begin)) {
    const Col16 C = col16_of(P, L);
    if (L.finest)
      k_gs32_range16<true, true><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.vc, L.col, L.val16,
                                                      L.diag32, r, x);
    else if (L.stream)
      k_gs32_range16<true, false><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.vc, L.col, L.val16,
                                                       L.diag32, r, x);
    else
      k_gs32_range16<false, false><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.vc, L.col, L.val16,
                                                        L.diag32, r, x);
    HPCGX_LAUNCHED();
    return;
  }
  if (L.finest)
    k_gs32_range<true, true><<<g, kBlock, 0, s>>>(begin, end, L.n, L.col,
                                                  L.val16, L.diag32, r, x);
  else if (L.stream)
    k_gs32_range<true, false><<<g, kBlock, 0, s>>>(begin, end, L.n, L.col,
                                                   L.val16, L.diag32, r, x);
  else
    k_gs32_range<false, false><<<g, kBlock, 0, s>>>(begin, end, L.n, L.col,
                                                    L.val16, L.diag32, r, x);
  HPCGX_LAUNCHED();
}

// gs32_step on a forward sweep that began from x = 0 (k_gs32_zero); false
// (nothing launched) where it does not apply.
bool gs32_step_zero(const Problem &P, const Level &L, int64_t begin, int64_t end,
                    const float *r, float *x, hipStream_t s) {
  if (!L.zmask || !L.vc || !use_col16(P, L, begin) || L.n >= ((int64_t)1 << 29)) return false;
  if (end <= begin) return true;
  const unsigned g = grid_for(end - begin);
  const Col16 C = col16_of(P, L);
  const uint32_t *vp = (L.v16p && P.use_pairs && P.use_c8) ? L.v16p : nullptr;
#define HPCGX_GZ32(ST, FI, LI)                                                                \
  k_gs32_zero<ST, FI, LI><<<g, kBlock, 0, s>>>(begin, end, L.n, C, L.zmask, vp, L.vc, L.col,    \
                                               L.val16, L.diag32, r, x, L.xcd)
  const bool live = P.zero_live & 4;
  if (L.finest && live) HPCGX_GZ32(true, true, true);
  else if (L.finest) HPCGX_GZ32(true, true, false);
  else if (L.stream && live) HPCGX_GZ32(true, false, true);
  else if (L.stream) HPCGX_GZ32(true, false, false);
  else if (live) HPCGX_GZ32(false, false, true);
  else HPCGX_GZ32(false, false, false);
#undef HPCGX_GZ32
  HPCGX_LAUNCHED();
  return true;
}

// Whether level l's first forward FP32 sweep of a V-cycle can start from an
// unfilled x (every colour step through gs32_step_zero).
bool zero_guess_sweep32(const Problem &P, int l, int ordering) {
  const Level &L = P.L[l];
  if (!P.zero_guess || ordering != HPCGX_ORDER_MULTICOLOR || !L.zmask || !L.vc ||
      !use_col16(P, L, 0) || L.n >= ((int64_t)1 << 29))
    return false;
  for (int c = 0; c < L.ncolors; ++c)
    if (L.color_ptr[c] % kTile) return false;
  return true;
}
}  // namespace

void prepare_mixed(Problem &P) {
  if (P.mixed_ready) return;
  TempBufD overflow(sizeof(int));
  HPCGX_HIP(hipMemset(overflow.p, 0, sizeof(int)));
  for (int l = 0; l < P.nlevels; ++l) {
    Level &L = P.L[l];
    L.val16 = reinterpret_cast<half_t *>(P.arena.alloc<uint16_t>((size_t)kSlots * L.n));
    L.diag32 = P.arena.alloc<float>(L.n);
    k_make_half<<<grid_for(L.n), kBlock>>>(L.n, L.val, L.dslot, L.val16, L.diag32,
                                           static_cast<int *>(overflow.p));
    HPCGX_LAUNCHED();
    if (L.c16) {
      const int64_t m = (int64_t)kSlots * L.n;
      L.vc = P.arena.alloc<uint32_t>((size_t)m);
      k_pack_vc<<<grid_for(L.n), kBlock>>>(L.n, L.col16, L.val16, L.vc);
      HPCGX_LAUNCHED();
      if (L.byte_tiles > 0) {
        L.v16p = P.arena.alloc<uint32_t>((size_t)kPairWords * L.n);
        k_pack_v16p<<<grid_for(L.n), kBlock>>>(L.n, L.val16, L.v16p);
        HPCGX_LAUNCHED();
      }
    }
    if (L.nc) {
      L.rc32 = P.arena.alloc<float>(L.nc);
      L.xc32 = P.arena.alloc<float>(L.nc);
    }
  }
  P.r32 = P.arena.alloc<float>(P.L[0].n);
  P.z32 = P.arena.alloc<float>(P.L[0].n);
  int bad = 0;
  HPCGX_HIP(hipMemcpy(&bad, overflow.p, sizeof bad, hipMemcpyDeviceToHost));
  if (bad & 1) throw Error(HPCGX_EUNSUP, "an off-diagonal value overflows FP16");
  P.half_exact = (bad & 2) == 0;
  P.mixed_ready = true;
}

void replace_diag32(Level &L, const double *d, hipStream_t s) {
  if (!L.diag32) return;
  k_diag32<<<grid_for(L.n), kBlock, 0, s>>>(L.n, d, L.diag32);
  HPCGX_LAUNCHED();
}

void symgs32(Problem &P, int l, int ordering, const float *r, float *x,
             hipStream_t s, bool zero_x, double *rz, const double *r64) {
  const Level &L = P.L[l];
  if (zero_x && !zero_guess_sweep32(P, l, ordering))
    throw Error(HPCGX_EINVAL, "zero-x FP32 sweep not available on this level");
  if (ordering == HPCGX_ORDER_MULTICOLOR) {
    for (int c = 0; c < L.ncolors; ++c) {
      if (zero_x)
        (void)gs32_step_zero(P, L, L.color_ptr[c], L.color_ptr[c + 1], r, x, s);
      else
        gs32_step(P, L, L.color_ptr[c], L.color_ptr[c + 1], r, x, s);
    }
    int64_t off = 0;
    for (int c = L.ncolors - 1; c >= 0; --c) {
      gs32_step(P, L, L.color_ptr[c], L.color_ptr[c + 1], r, x, s, rz ? rz + off : nullptr,
                r64);
      off += rz_slots(L.color_ptr[c + 1] - L.color_ptr[c]);
    }
  } else {
    for (int pass = 0; pass < 2; ++pass)
      for (int k = 0; k < L.nsched; ++k) {
        const int t = pass == 0 ? k : L.nsched - 1 - k;
        const int64_t cnt = L.level_ptr[t + 1] - L.level_ptr[t];
        k_gs32_list<<<grid_for(cnt), kBlock, 0, s>>>(
            L.level_rows + L.level_ptr[t], cnt, L.n, L.col, L.val16, L.diag32,
            r, x);
        HPCGX_LAUNCHED();
      }
  }
}

void spmv32(const Problem &P, const Level &L, const float *x, float *y, hipStream_t s) {
  if (use_col16(P, L, 0)) {
    const Col16 C = col16_of(P, L);
    if (L.stream)
      k_spmv32_16<true><<<(unsigned)L.nchunks, kBlock, 0, s>>>(L.chunks, L.n, C, L.vc,
                                                               L.col, L.val16, L.diag32, x, y);
    else
      k_spmv32_16<false><<<(unsigned)L.nchunks, kBlock, 0, s>>>(L.chunks, L.n, C, L.vc,
                                                                L.col, L.val16, L.diag32, x, y);
    HPCGX_LAUNCHED();
    return;
  }
  if (L.stream)
    k_spmv32<true><<<(unsigned)L.nchunks, kBlock, 0, s>>>(L.chunks, L.n, L.col,
                                                          L.val16, L.diag32, x, y);
  else
    k_spmv32<false><<<(unsigned)L.nchunks, kBlock, 0, s>>>(L.chunks, L.n, L.col,
                                                           L.val16, L.diag32, x, y);
  HPCGX_LAUNCHED();
}

// ComputeMG_ref.cpp:39-63 in FP32.  x = 0 (:42) is a fill kernel, or,
// where zero_guess_sweep32 allows, implied by the first forward sweep.
void mg32(Problem &P, int l, int ordering, const float *r, float *x,
          hipStream_t s, double *rz, const double *r64) {
  Level &L = P.L[l];
  const bool zg = zero_guess_sweep32(P, l, ordering);
  if (!zg) {
    k_fill32<<<grid_for(L.n), kBlock, 0, s>>>(x, L.n);
    HPCGX_LAUNCHED();
  }
  if (rz && l + 1 >= P.nlevels)
    throw Error(HPCGX_EINVAL, "r.z partials come from a post-smoother sweep");
  symgs32(P, l, ordering, r, x, s, zg);
  if (l + 1 < P.nlevels) {
    const unsigned g = grid_for(L.nc);
    if (L.rperm) {
      const Col16 C = col16_of(P, L);
      const bool c16 = L.vc && use_col16(P, L, 0);
      const bool pairs = c16 && L.v16p && P.use_pairs && P.use_c8;
#define HPCGX_RR32(ST, K)                                                                  \
  k_resid_restrict32_f<ST, K><<<g, kBlock, 0, s>>>(L.nc, L.ftile, L.rperm, L.n, C, L.vc, L.v16p, \
                                                   L.col, L.val16, L.diag32, r, x, L.rc32)
      if (L.stream && pairs) HPCGX_RR32(true, 2);
      else if (L.stream && c16) HPCGX_RR32(true, 1);
      else if (L.stream) HPCGX_RR32(true, 0);
      else if (pairs) HPCGX_RR32(false, 2);
      else if (c16) HPCGX_RR32(false, 1);
      else HPCGX_RR32(false, 0);
#undef HPCGX_RR32
    } else if (L.stream) {
      k_resid_restrict32<true><<<g, kBlock, 0, s>>>(L.nc, L.n, L.f2c, L.col, L.val16,
                                                    L.diag32, r, x, L.rc32);
    } else {
      k_resid_restrict32<false><<<g, kBlock, 0, s>>>(L.nc, L.n, L.f2c, L.col, L.val16,
                                                     L.diag32, r, x, L.rc32);
    }
    HPCGX_LAUNCHED();
    mg32(P, l + 1, ordering, L.rc32, L.xc32, s);
    if (L.rperm)
      k_prolong32_f<<<g, kBlock, 0, s>>>(L.nc, L.ftile, L.rperm, L.xc32, x);
    else
      k_prolong32<<<g, kBlock, 0, s>>>(L.nc, L.f2c, L.xc32, x);
    HPCGX_LAUNCHED();
    symgs32(P, l, ordering, r, x, s, false, rz, r64);
  }
}

void to_f32(int64_t n, const double *in, float *out, hipStream_t s) {
  k_to_f32<<<grid_for(n), kBlock, 0, s>>>(n, in, out);
  HPCGX_LAUNCHED();
}

void to_f64(int64_t n, const float *in, double *out, hipStream_t s) {
  k_to_f64<<<grid_for(n), kBlock, 0, s>>>(n, in, out);
  HPCGX_LAUNCHED();
}

namespace {
double read_scalar(const Problem &P, int slot, hipStream_t s) {
  double v = 0.0;
  HPCGX_HIP(hipMemcpyAsync(&v, P.scalars + slot, sizeof v,
                           hipMemcpyDeviceToHost, s));
  HPCGX_HIP(hipStreamSynchronize(s));
  return v;
}
}  // namespace

// CG.cpp:59-143 with the FP32 V-cycle as M^-1 and the FP64 residual
// recomputed every `recompute` iterations.
namespace {
// One iteration of the mixed-precision CG without the residual recompute
// (CG.cpp:102-124 with the FP32 V-cycle), enqueued on s.
// r.z partials the finest level's last FP32 colour sweep leaves (one per
// wavefront of rows), or 0 where that sweep cannot leave them.
int64_t rz_partials32(const Problem &P, int ordering) {
  const Level &L = P.L[0];
  if (!P.rz_fused || !P.rz_part || ordering != HPCGX_ORDER_MULTICOLOR || P.nlevels < 2)
    return 0;
  int64_t count = 0;
  for (int c = 0; c < L.ncolors; ++c) {
    const int64_t b = L.color_ptr[c], e = L.color_ptr[c + 1];
    if (e <= b) continue;
    if (!pairs_step(P, L, b)) return 0;
    count += rz_slots(e - b);
  }
  return count <= P.nrz_partials ? count : 0;
}

void mixed_iteration(Problem &P, int ordering, double *x, bool first, hipStream_t s) {
  const Level &A = P.L[0];
  const int64_t n = A.n;
  const int64_t nrz = rz_partials32(P, ordering);
  mg32(P, 0, ordering, P.r32, P.z32, s, nrz ? P.rz_part : nullptr, P.r);
  if (nrz)  // r.z from the V-cycle's last sweep (CG.cpp:110, :113-114)
    launch_partials_dot(P, P.rz_part, nrz, kFinishBeta, s);
  else
    launch_dot32(P, n, P.r, P.z32, nullptr, kFinishBeta, s);
  if (first)
    to_f64(n, P.z32, P.p, s);
  else if (P.defer_x)  // with the previous iteration's x step (defer_x)
    launch_update_xp32(n, x, P.z32, P.p, P.scalars, s);
  else
    launch_update_p32(n, P.z32, P.p, P.scalars, s, P.stream_p);
  launch_spmv_dot_half(P, A, P.p, P.Ap, s);
  launch_update_xr(P, n, P.defer_x ? nullptr : x, P.r, P.p, P.Ap, s, P.r32);
}

// The captured body of iterations 2.. (rebuilt when its key changes; fills
// are kernels, so the graph holds no memset nodes).
hipGraphExec_t mixed_graph(Problem &P, int ordering, double *x) {
  auto &G = P.mgraph;
  if (G.exec && G.x == x && G.ordering == ordering) return G.exec;
  G.reset();
  if (!G.cap) HPCGX_HIP(hipStreamCreateWithFlags(&G.cap, hipStreamNonBlocking));
  HPCGX_HIP(hipStreamBeginCapture(G.cap, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    mixed_iteration(P, ordering, x, false, G.cap);
  } catch (...) {
    (void)hipStreamEndCapture(G.cap, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HPCGX_HIP(hipStreamEndCapture(G.cap, &graph));
  const hipError_t e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HPCGX_HIP(e);
  G.x = x;
  G.ordering = ordering;
  return G.exec;
}
}  // namespace

void cg_mixed(Problem &P, int ordering, const double *b, double *x,
              int max_iter, double tol, int recompute, int *niters,
              double *normr, double *normr0, double *times, hipStream_t s) {
  prepare_mixed(P);
  const auto t_begin = std::chrono::steady_clock::now();
  const Level &A = P.L[0];
  const int64_t n = A.n;
  *niters = 0;
  launch_copy(x, P.p, n, s);
  launch_spmv(P, A, P.p, P.Ap, s);
  launch_residual0(P, n, b, P.Ap, P.r, s);
  double nr = std::sqrt(read_scalar(P, kNormr2, s));
  *normr0 = nr;
  const bool check_each = tol > 0.0;
  // r32 = r rounded to FP32 is kept up to date by the r update itself; z
  // stays FP32 (z32) and is widened inside the dot and the p update, which
  // gives the bits of widening it first.
  to_f32(n, P.r, P.r32, s);
  for (int k = 1; k <= max_iter && nr / *normr0 > tol; ++k) {
    if (k >= 2 && P.use_mixed_graph && !P.klog.on)
      HPCGX_HIP(hipGraphLaunch(mixed_graph(P, ordering, x), s));
    else
      mixed_iteration(P, ordering, x, k == 1, s);
    if (recompute > 0 && k % recompute == 0) {
      if (P.defer_x) {
        // x's deferred step now; alpha then reads 0 in the next p update,
        // whose x step (x + 0 p) leaves x as it is
        launch_step_x(n, x, P.p, P.scalars, s);
        launch_fill(P.scalars + kAlpha, 1, 0.0, s);
      }
      // r = b - A x in FP64 (replaces the recursively updated r)
      if (!launch_residual_fused(P, A, b, x, P.r, P.r32, s)) {
        launch_spmv(P, A, x, P.Ap, s);
        launch_residual0(P, n, b, P.Ap, P.r, s);
        to_f32(n, P.r, P.r32, s);
      }
    }
    *niters = k;
    if (check_each || k == max_iter) nr = std::sqrt(read_scalar(P, kNormr2, s));
  }
  // the last iteration's deferred x step (none after a recompute, which took
  // it, or before the first iteration)
  if (P.defer_x && *niters > 0 && !(recompute > 0 && *niters % recompute == 0))
    launch_step_x(n, x, P.p, P.scalars, s);
  if (!check_each) nr = std::sqrt(read_scalar(P, kNormr2, s));
  HPCGX_HIP(hipStreamSynchronize(s));
  *normr = nr;
  if (times)
    times[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                              t_begin).count();
  if (!std::isfinite(nr)) throw Error(HPCGX_EINVAL, "CG residual is not finite");
}

}  // namespace hpcgx
