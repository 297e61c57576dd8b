// setup.hip -- device-side problem setup: generation, schedules, reordering.
//
// The schedule builder is generic (it reads only the sparsity structure):
//  * the lexicographic level schedule is the level-set partition of the
//    dependency DAG "row i needs every column j < i of row i", found with a
//    level-synchronous Kahn traversal: a row becomes ready when its last
//    lower neighbour has been placed, and its level is one more than that
//    neighbour's;
//  * the colour of a row is the smallest colour not used by its lower
//    neighbours, assigned the moment the row becomes ready.  That is exactly
//    the greedy first-fit colouring of the reference's OptimizeProblem.cpp:
//    50-79 (which visits rows in natural order and looks only at columns
//    j < i), computed in parallel because a row's colour depends only on
//    rows of lower levels.
// Rows are then renumbered colour by colour, in the colours' sweep order
// (kSweepOrders; natural order inside a colour), and the 27-slot ELL is
// rebuilt in that order, with the slots of every tile of 64 rows aligned by
// natural column offset so that the column ids of a slot fit 16 bits above a
// per-tile base (k_tile_layout).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <numeric>

#include <climits>

#include "internal.hpp"

namespace hpcgx {
namespace {

// Forward visiting orders of the greedy colours 0..7 (HPCGX_SWEEP_ORDER);
// colours from 8 on follow in id order.  On the 27-point grid colour
// (x&1) + 2(y&1) + 4(z&1) is the parity class, and colour 0 holds the points
// the restriction injects (ComputeRestriction_ref.cpp:51 through f2c,
// GenerateCoarseProblem.cpp:77-79).  Swept last, as in id order, colour 0's
// rows end each backward sweep with a residual of zero up to rounding, so
// the coarse grid receives almost nothing.  Order 0 visits colour 0 last in
// the forward sweep and first in the backward one; at 512x512x640 PCG then
// reaches the 50-iteration lexicographic residual in 57 iterations instead
// of 70.  Order 1 is the id order.
constexpr int kSweepOrders[][8] = {
    {3, 5, 6, 7, 1, 2, 4, 0},
    {0, 1, 2, 3, 4, 5, 6, 7},
};
constexpr int kNumSweepOrders = (int)(sizeof(kSweepOrders) / sizeof(kSweepOrders[0]));

constexpr int kStepBlocks = 2048;      // fixed grid of the Kahn steps
constexpr int kSyncEvery = 128;        // Kahn steps between host checks
constexpr int64_t kMaxSched = 1 << 22; // bound on schedule levels

inline unsigned grid_for(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}
__device__ inline int64_t gtid() {
  return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}

struct TempBuf {  // a temporary device buffer freed on scope exit
  void *p = nullptr;
  explicit TempBuf(size_t bytes) { HPCGX_HIP(hipMalloc(&p, bytes ? bytes : 1)); }
  ~TempBuf() { (void)hipFree(p); }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// GenerateProblem_ref.cpp:131-183 (single rank): natural order, neighbours
// in sz, sy, sx order, 26 on the diagonal and -1 elsewhere.  The grid's rows
// are rows [off, off + nx*ny*nz) of a level of ntot rows (a fleet level
// holds several grids one after the other; a single problem has off = 0).
__global__ __launch_bounds__(kBlock) void k_generate(int nx, int ny, int nz,
                                                      int64_t n, int64_t off,
                                                      int64_t ntot, int32_t *col,
                                                      double *val,
                                                      uint8_t *nnz_row) {
  const int64_t i = gtid();
  if (i >= n) return;
  const int ix = (int)(i % nx);
  const int64_t t = i / nx;
  const int iy = (int)(t % ny);
  const int iz = (int)(t / ny);
  const int64_t nxy = (int64_t)nx * ny;
  const int64_t g = off + i;
  int cnt = 0;
  for (int sz = -1; sz <= 1; ++sz) {
    if (iz + sz < 0 || iz + sz >= nz) continue;
    for (int sy = -1; sy <= 1; ++sy) {
      if (iy + sy < 0 || iy + sy >= ny) continue;
      for (int sx = -1; sx <= 1; ++sx) {
        if (ix + sx < 0 || ix + sx >= nx) continue;
        const int64_t c = g + sz * nxy + (int64_t)sy * nx + sx;
        col[(int64_t)cnt * ntot + g] = (int32_t)c;
        val[(int64_t)cnt * ntot + g] = (c == g) ? 26.0 : -1.0;
        ++cnt;
      }
    }
  }
  for (int j = cnt; j < kSlots; ++j) {
    col[(int64_t)j * ntot + g] = (int32_t)g;
    val[(int64_t)j * ntot + g] = 0.0;
  }
  nnz_row[g] = (uint8_t)cnt;
}

// GenerateCoarseProblem.cpp:71-82, for a grid whose fine rows start at offf
// and whose coarse rows start at offc.
__global__ __launch_bounds__(kBlock) void k_f2c(int nxf, int nyf, int64_t nc,
                                                 int nxc, int nyc, int64_t offf,
                                                 int64_t offc, int32_t *f2c) {
  const int64_t c = gtid();
  if (c >= nc) return;
  const int ixc = (int)(c % nxc);
  const int64_t t = c / nxc;
  const int iyc = (int)(t % nyc);
  const int izc = (int)(t / nyc);
  f2c[offc + c] = (int32_t)(offf + (int64_t)(2 * izc) * nxf * nyf +
                            (int64_t)(2 * iyc) * nxf + 2 * ixc);
}

__global__ __launch_bounds__(kBlock) void k_count_nnz(int64_t n,
                                                       const uint8_t *nnz_row,
                                                       unsigned long long *out) {
  const int64_t i = gtid();
  unsigned long long v = (i < n) ? nnz_row[i] : 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, v);
}

// Structural checks: columns in range, a diagonal entry, symmetric pattern
// (the Kahn traversal walks each row's upper entries as the rows that
// depend on it).  bad: 1 = column out of range, 2 = no diagonal,
// 4 = pattern not symmetric.
__global__ __launch_bounds__(kBlock) void k_check(int64_t n, const int32_t *col,
                                                   const uint8_t *nnz_row,
                                                   int32_t *lower_count,
                                                   int *bad,
                                                   unsigned long long *bandwidth) {
  const int64_t i = gtid();
  if (i >= n) return;
  const int m = nnz_row[i];
  int lower = 0, has_diag = 0, flags = 0;
  int64_t bw = 0;
  for (int j = 0; j < m; ++j) {
    const int64_t c = col[(int64_t)j * n + i];
    if (c < 0 || c >= n) { flags |= 1; continue; }
    if (c == i) { has_diag = 1; continue; }
    bw = max(bw, c > i ? c - i : i - c);
    if (c < i) ++lower;
    const int mc = nnz_row[c];
    int found = 0;
    for (int k = 0; k < mc; ++k) found |= (col[(int64_t)k * n + c] == i);
    if (!found) flags |= 4;
  }
  if (!has_diag) flags |= 2;
  lower_count[i] = lower;
  if (flags) atomicOr(bad, flags);
  unsigned long long w = (unsigned long long)bw;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_down(w, off, 64);
    w = o > w ? o : w;
  }
  if ((threadIdx.x & 63) == 0 && w) atomicMax(bandwidth, w);
}

// Append `want` lanes of the calling wavefront to a list with one atomic per
// wavefront (all active lanes must call it together).  Returns the slot of
// the lane, or -1.
__device__ inline int64_t wave_append(bool want, unsigned *tail) {
  const unsigned long long mask = __ballot(want);
  if (mask == 0ull) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(tail, (unsigned)__popcll(mask));
  base = __shfl(base, leader, 64);
  const unsigned rank = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
  return want ? (int64_t)(base + rank) : -1;
}

// Level 0 of the traversal: rows without lower neighbours.
__global__ __launch_bounds__(kBlock) void k_kahn_init(int64_t n,
                                                       const int32_t *remaining,
                                                       int32_t *order,
                                                       unsigned *tail) {
  const int64_t i = gtid();
  const bool ready = i < n && remaining[i] == 0;
  const int64_t slot = wave_append(ready, tail);
  if (ready) order[slot] = (int32_t)i;
}

// Step t: the rows of level t are order[lptr[t] .. lptr[t+1]).  Each row is
// handled by 32 lanes, one per slot:
//  * its colour is the smallest one absent from its lower neighbours (all on
//    lower levels, coloured in earlier steps): the lanes OR their
//    neighbour's colour bit together;
//  * it releases its upper neighbours: a neighbour whose last lower
//    neighbour this was is appended to `order` as a row of level t+1.
// Unused slots hold the row itself and do neither.
__global__ __launch_bounds__(kBlock) void k_kahn_step(
    int t, const int64_t *lptr, int64_t n, const int32_t *col,
    int32_t *remaining, int32_t *level, int32_t *color, int32_t *order,
    unsigned *tail) {
  const int64_t begin = lptr[t], end = lptr[t + 1];
  const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 5;
  const int j = threadIdx.x & 31;
  for (int64_t k = begin + (gtid() >> 5); k < end; k += stride) {
    const int64_t u = order[k];
    const int64_t c = (j < kSlots) ? col[(int64_t)j * n + u] : u;
    unsigned bit = (c < u) ? (1u << color[c]) : 0u;
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) bit |= __shfl_xor(bit, off, 32);
    if (j == 0) {
      level[u] = t;
      color[u] = __ffs(~bit) - 1;
    }
    bool ready = false;
    if (c > u) ready = atomicSub(&remaining[c], 1) == 1;
    const int64_t slot = wave_append(ready, tail);
    if (ready) order[slot] = (int32_t)c;
  }
}

// Proposed schedule for rows that come from known grids: level x + 2y + 4z
// and colour (x & 1) + 2 (y & 1) + 4 (z & 1) of the row's own grid.  A row
// outside every part gets -1 (the verification then rejects the proposal).
__global__ __launch_bounds__(kBlock) void k_propose_schedule(int64_t n, int nparts,
                                                             const GridPart *__restrict__ parts,
                                                             int32_t *level, int32_t *color) {
  const int64_t i = gtid();
  if (i >= n) return;
  int lo = 0, hi = nparts - 1;  // last part with off <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (parts[mid].off <= i) lo = mid; else hi = mid - 1;
  }
  const GridPart p = parts[lo];
  const int64_t local = i - p.off;
  if (local < 0 || local >= (int64_t)p.nx * p.ny * p.nz) {
    level[i] = -1;
    color[i] = -1;
    return;
  }
  const int x = (int)(local % p.nx);
  const int64_t t = local / p.nx;
  const int y = (int)(t % p.ny), z = (int)(t / p.ny);
  level[i] = x + 2 * y + 4 * z;
  color[i] = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
}

// Accepts a proposed (level, colour) only if every row has exactly what the
// Kahn traversal gives it: level = 1 + the largest level of its lower
// neighbours (0 without any), colour = the smallest colour none of its lower
// neighbours has.  By induction over the rows in natural order the proposal
// then equals the traversal's result.  bad: 1 = a row disagrees.
__global__ __launch_bounds__(kBlock) void k_verify_schedule(int64_t n, const int32_t *col,
                                                            const uint8_t *nnz_row,
                                                            const int32_t *level,
                                                            const int32_t *color, int *bad,
                                                            int *max_level) {
  const int64_t i = gtid();
  bool ok = true;
  int lv = 0;
  if (i < n) {
    const int m = nnz_row[i];
    unsigned used = 0;
    int below = -1;
    for (int j = 0; j < m; ++j) {
      const int64_t c = col[(int64_t)j * n + i];
      if (c >= i) continue;
      const int cc = color[c];
      ok = ok && cc >= 0 && cc < 32;
      if (cc >= 0 && cc < 32) used |= 1u << cc;
      below = max(below, (int)level[c]);
    }
    lv = level[i];
    ok = ok && color[i] == __ffs(~used) - 1 && lv == below + 1;
  }
  if (!ok) atomicOr(bad, 1);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) lv = max(lv, __shfl_xor(lv, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(max_level, lv);
}

__global__ void k_kahn_advance(int t, int64_t *lptr, const unsigned *tail) {
  lptr[t + 2] = (int64_t)*tail;
}

__global__ __launch_bounds__(kBlock) void k_fill_i64(int64_t *a, int64_t n,
                                                      int64_t v) {
  const int64_t i = gtid();
  if (i < n) a[i] = v;
}

// first[key] = first position of key in a sorted key array.
__global__ __launch_bounds__(kBlock) void k_bounds(int64_t n,
                                                    const uint32_t *sorted,
                                                    int64_t *first) {
  const int64_t i = gtid();
  if (i < n && (i == 0 || sorted[i - 1] != sorted[i])) first[sorted[i]] = i;
}

__global__ __launch_bounds__(kBlock) void k_iota(int64_t n, int32_t *a) {
  const int64_t i = gtid();
  if (i < n) a[i] = (int32_t)i;
}

__global__ __launch_bounds__(kBlock) void k_invert(int64_t n,
                                                    const int32_t *iperm,
                                                    int32_t *perm) {
  const int64_t k = gtid();
  if (k < n) perm[iperm[k]] = (int32_t)k;
}

__global__ __launch_bounds__(kBlock) void k_map(int64_t n, const int32_t *map,
                                                 const int32_t *in,
                                                 int32_t *out) {
  const int64_t k = gtid();
  if (k < n) out[k] = map[in[k]];
}

__device__ inline int wave_min_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int wave_max_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Tile-aligned slots and 16-bit column ids (Level::col16), one wavefront per
// tile of kTile device rows, in place on the device-order ELL.
//  1. The tile's distinct natural column offsets (column - row, natural
//     numbering), ascending: repeated wavefront minimum of each row's next
//     offset.  A row's offsets ascend with its columns, so entry e of a row
//     lands in a slot >= e and the row's entries keep their order.
//  2. Entries move to their slots (highest slot first, so nothing is
//     overwritten before it is read); a row without an entry in a slot gets
//     a column already in that slot's range and the value 0.  The slot's base
//     is its smallest column.  The tile's class: 8-bit offsets if every slot
//     spans at most 255 columns, 16-bit if at most 65535, else (or with more
//     than 27 distinct offsets) 32-bit ids only.
__global__ __launch_bounds__(kBlock) void k_tile_layout(
    int64_t n, int64_t ntiles, const int32_t *__restrict__ iperm, int32_t *col,
    double *val, const uint8_t *__restrict__ nnz_row, uint8_t *dslot,
    uint16_t *col16, uint32_t *off8, int32_t *cbase, int32_t *tbase, int32_t *tinfo,
    uint8_t *amask) {
  const int64_t t = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (t >= ntiles) return;  // whole wavefronts
  const int lane = threadIdx.x & 63;
  const int64_t i = t * kTile + lane;
  const bool row = i < n;
  const int m = row ? nnz_row[i] : 0;
  const int64_t inat = row ? iperm[i] : 0;

  unsigned mask = 0;  // slots holding one of this row's entries
  int ptr = 0, nu = 0;
  bool fits = true;
  for (;;) {
    int head = INT_MAX;
    if (ptr < m) head = (int)((int64_t)iperm[col[(int64_t)ptr * n + i]] - inat);
    const int mn = wave_min_i32(head);
    if (mn == INT_MAX) break;
    if (nu == kSlots) {
      fits = false;
      break;
    }
    if (head == mn) {
      mask |= 1u << nu;
      ++ptr;
    }
    ++nu;
  }
  if (!fits) {
    if (lane == 0) {
      tinfo[t] = kTile32;
      amask[t] = 0;
    }
    return;
  }

  bool narrow = true, narrow8 = true;
  unsigned affine = 0;  // slots whose column is base + lane in every row
  int ds = 0;
  uint32_t o8[kSlotWords] = {};
#pragma unroll
  for (int u = kSlots - 1; u >= 0; --u) {
    const bool has = (mask >> u) & 1u;
    int32_t c = 0;
    double v = 0.0;
    if (has) {
      const int e = __popc(mask & ((1u << u) - 1u));
      c = col[(int64_t)e * n + i];
      v = val[(int64_t)e * n + i];
    }
    const int lo = wave_min_i32(has ? c : INT_MAX);
    const int hi = wave_max_i32(has ? c : INT_MIN);
    const bool used = lo != INT_MAX;
    if (used && hi - lo > 65535) narrow = false;
    if (used && hi - lo > 255) narrow8 = false;
    const int32_t base = used ? lo : (int32_t)(t * kTile);
    if (row) {
      const int32_t cc = has ? c : (used ? lo : (int32_t)i);
      col[(int64_t)u * n + i] = cc;
      val[(int64_t)u * n + i] = has ? v : 0.0;
      col16[(int64_t)u * n + i] = (uint16_t)(cc - base);
      o8[u / 4] |= ((uint32_t)(cc - base) & 0xffu) << (8 * (u % 4));
      if (has && c == i) ds = u;
    }
    const int32_t cl = row ? (has ? c : (used ? lo : (int32_t)i)) : base + lane;
    if (wave_min_i32(cl - base == lane ? 1 : 0) == 1) affine |= 1u << u;
    if (lane == 0) {
      cbase[(int64_t)u * ntiles + t] = base;
      tbase[t * kBaseStride + u] = base;
    }
  }
  if (row) {
    dslot[i] = (uint8_t)ds;
#pragma unroll
    for (int g = 0; g < kSlotWords; ++g) off8[(int64_t)g * n + i] = o8[g];
  }
  if (lane == 0) {
    tinfo[t] = !narrow ? kTile32 : narrow8 ? kTile8 : kTile16;
    uint8_t m = 0;  // offset word g: slots 4g .. 4g+3 (below kSlots) all affine
    for (int g = 0; g < kSlotWords; ++g) {
      const unsigned want = ((1u << 4) - 1u) << (4 * g) & ((1u << kSlots) - 1u);
      if ((affine & want) == want) m |= (uint8_t)(1u << g);
    }
    amask[t] = narrow8 ? m : 0;
  }
}

// Rebuild the ELL in device order: device row k is natural row iperm[k];
// columns are renumbered through perm; unused slots become (k, 0.0).
__global__ __launch_bounds__(kBlock) void k_permute_ell(
    int64_t n, const int32_t *__restrict__ iperm,
    const int32_t *__restrict__ perm, const int32_t *__restrict__ col_nat,
    const double *__restrict__ val_nat, const uint8_t *__restrict__ nnz_nat,
    int32_t *__restrict__ col, double *__restrict__ val,
    double *__restrict__ diag, uint8_t *__restrict__ dslot,
    uint8_t *__restrict__ nnz_row) {
  const int64_t k = gtid();
  if (k >= n) return;
  const int64_t i = iperm[k];
  const int m = nnz_nat[i];
  int ds = -1;
  for (int j = 0; j < kSlots; ++j) {
    const int64_t src = (int64_t)j * n + i, dst = (int64_t)j * n + k;
    if (j < m) {
      const int32_t c = col_nat[src];
      if (c == i && ds < 0) ds = j;
      col[dst] = perm[c];
      val[dst] = val_nat[src];
    } else {
      col[dst] = (int32_t)k;
      val[dst] = 0.0;
    }
  }
  dslot[k] = (uint8_t)ds;
  diag[k] = val_nat[(int64_t)ds * n + i];
  nnz_row[k] = (uint8_t)m;
}

// f2c in device order: coarse device row kc -> fine device row.
__global__ __launch_bounds__(kBlock) void k_link(int64_t nc,
                                                  const int32_t *iperm_c,
                                                  const int32_t *f2c_nat,
                                                  const int32_t *perm_f,
                                                  int32_t *f2c) {
  const int64_t kc = gtid();
  if (kc < nc) f2c[kc] = perm_f[f2c_nat[iperm_c[kc]]];
}

__global__ __launch_bounds__(kBlock) void k_vectors(int64_t n,
                                                     const uint8_t *nnz_row,
                                                     double *b, double *x,
                                                     double *xexact) {
  const int64_t k = gtid();
  if (k >= n) return;
  if (b) b[k] = 26.0 - ((double)(nnz_row[k] - 1));  // GenerateProblem_ref.cpp:178
  if (x) x[k] = 0.0;
  if (xexact) xexact[k] = 1.0;
}

int bits_for(int64_t v) {
  int b = 1;
  while (((int64_t)1 << b) <= v) ++b;
  return b;
}

// Device-order key: colour, then the band of `band` grid lines (y) the
// natural row lies in.  Inside a colour, rows then run band by band, each
// band plane by plane: the x values a colour step gathers from the
// neighbouring planes are reused one band later instead of one whole plane
// later, which keeps them in the L2 cache.
__global__ __launch_bounds__(kBlock) void k_band_key(int64_t n, int nx, int ny, int band,
                                                     int nbands, const int32_t *color,
                                                     int32_t *key) {
  const int64_t i = gtid();
  if (i >= n) return;
  const int y = (int)((i / nx) % ny);
  key[i] = color[i] * nbands + y / band;
}

// Colour visiting order of the multicolour sweeps: colour c (a greedy
// first-fit id) is visited in position rank[c] of the forward sweep (the
// backward sweep visits the reverse).
struct ColourRank {
  int r[kMaxColors];
};
__global__ __launch_bounds__(kBlock) void k_colour_rank(int64_t n, const int32_t *color,
                                                        ColourRank rank, int32_t *key) {
  const int64_t i = gtid();
  if (i >= n) return;
  key[i] = rank.r[color[i]];
}

// Stable sort of row ids by a small non-negative integer key (< nbins):
// rows_out = rows ordered by (key, row); returns the bin pointers
// ptr[0..nbins], bin k owning rows_out[ptr[k] .. ptr[k+1]).
std::vector<int64_t> sort_rows_by_key(int64_t n, const int32_t *key,
                                      int nbins, int32_t *rows_out,
                                      hipStream_t s) {
  HPCGX_REQUIRE(n < (int64_t)1 << 31, "row count exceeds 2^31");
  TempBuf ids(sizeof(int32_t) * n), keys_out(sizeof(uint32_t) * n);
  k_iota<<<grid_for(n), kBlock, 0, s>>>(n, ids.as<int32_t>());
  HPCGX_LAUNCHED();
  const uint32_t *ukey = reinterpret_cast<const uint32_t *>(key);
  const int end_bit = bits_for(nbins);
  size_t temp_bytes = 0;
  HPCGX_HIP(hipcub::DeviceRadixSort::SortPairs(
      nullptr, temp_bytes, ukey, keys_out.as<uint32_t>(), ids.as<int32_t>(),
      rows_out, (int)n, 0, end_bit, s));
  {
    TempBuf temp(temp_bytes);
    HPCGX_HIP(hipcub::DeviceRadixSort::SortPairs(
        temp.p, temp_bytes, ukey, keys_out.as<uint32_t>(), ids.as<int32_t>(),
        rows_out, (int)n, 0, end_bit, s));
    HPCGX_HIP(hipStreamSynchronize(s));
  }
  // bin starts from the sorted keys; empty bins start where the next begins
  TempBuf first(sizeof(int64_t) * (nbins + 1));
  k_fill_i64<<<grid_for(nbins + 1), kBlock, 0, s>>>(first.as<int64_t>(),
                                                    nbins + 1, n);
  HPCGX_LAUNCHED();
  k_bounds<<<grid_for(n), kBlock, 0, s>>>(n, keys_out.as<uint32_t>(),
                                          first.as<int64_t>());
  HPCGX_LAUNCHED();
  std::vector<int64_t> ptr(nbins + 1);
  HPCGX_HIP(hipMemcpyAsync(ptr.data(), first.p, sizeof(int64_t) * (nbins + 1),
                           hipMemcpyDeviceToHost, s));
  HPCGX_HIP(hipStreamSynchronize(s));
  ptr[nbins] = n;
  for (int b = nbins - 1; b >= 0; --b) ptr[b] = std::min(ptr[b], ptr[b + 1]);
  return ptr;
}

// tile_ptr[c*(S+1) + s] = first device row of colour c whose natural row is
// >= s*W (binary search: a colour block lists its natural rows in order).
__global__ __launch_bounds__(kBlock) void k_tiles(int C, int S, int64_t W,
                                                   const int64_t *color_ptr,
                                                   const int32_t *iperm,
                                                   int64_t *tile_ptr) {
  const int64_t k = gtid();
  if (k >= (int64_t)C * (S + 1)) return;
  const int c = (int)(k / (S + 1)), sl = (int)(k % (S + 1));
  int64_t lo = color_ptr[c], hi = color_ptr[c + 1];
  const int64_t key = (int64_t)sl * W;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if ((int64_t)iperm[mid] < key) lo = mid + 1; else hi = mid;
  }
  tile_ptr[k] = lo;
}

// Level::zmask: one wavefront per tile.  Row i's colour begins at the last
// colour pointer <= i; slot j is live for the tile when some row holds a
// nonzero there whose column lies below its colour's first row.
__global__ __launch_bounds__(kBlock) void k_zero_mask(int64_t n, int64_t T, int ncolors,
                                                       const int64_t *__restrict__ cptr,
                                                       const int32_t *__restrict__ col,
                                                       const double *__restrict__ val,
                                                       uint32_t *__restrict__ zmask) {
  const int64_t t = gtid() / 64;
  const int lane = threadIdx.x & 63;
  if (t >= T) return;  // uniform over the wavefront
  const int64_t i = t * kTile + lane;
  int64_t cb = 0;
  if (i < n) {
    int c = 0;
    while (c + 1 < ncolors && cptr[c + 1] <= i) ++c;
    cb = cptr[c];
  }
  uint32_t m = 0;
  for (int j = 0; j < kSlots; ++j) {
    bool live = false;
    if (i < n) {
      const int64_t k = (int64_t)j * n + i;
      live = val[k] != 0.0 && col[k] < cb;
    }
    if (__ballot(live)) m |= 1u << j;
  }
  if (lane == 0) zmask[t] = m;
}

}  // namespace

// Row chunks of at most kBlock rows.  Interleaved: ordered by their relative
// position in their colour block and then by colour (chunk k of every colour
// before chunk k+1 of any), so a sweep reads each colour block's x values
// at about the same relative position at the same time.  Otherwise colour
// block by colour block: the persistent SpMV's XCD x (which takes the x-th
// eighth of the list) then streams about one colour block's matrix rows in
// order, as a colour step does.
void build_chunks(Problem &P, Level &L, bool interleave) {
  if (L.chunks) P.arena.release(L.chunks, sizeof(int64_t) * 2 * L.nchunks);
  L.chunks = nullptr;
  struct Chunk {
    double key;
    int colour;
    int64_t begin, end;
  };
  std::vector<Chunk> v;
  for (int c = 0; c < L.ncolors; ++c) {
    const int64_t b0 = L.color_ptr[c], e0 = L.color_ptr[c + 1];
    const int64_t nch = (e0 - b0 + kBlock - 1) / kBlock;
    for (int64_t k = 0; k < nch; ++k)
      v.push_back({(k + 0.5) / (double)nch, c, b0 + k * kBlock,
                   std::min<int64_t>(b0 + (k + 1) * kBlock, e0)});
  }
  if (interleave)
    std::stable_sort(v.begin(), v.end(), [](const Chunk &a, const Chunk &b) {
      return a.key < b.key || (a.key == b.key && a.colour < b.colour);
    });
  L.chunks_interleaved = interleave;
  std::vector<int64_t> h(2 * v.size());
  for (size_t k = 0; k < v.size(); ++k) {
    h[2 * k] = v[k].begin;
    h[2 * k + 1] = v[k].end;
  }
  L.nchunks = (int64_t)v.size();
  L.chunks = P.arena.alloc<int64_t>(h.size());
  HPCGX_HIP(hipMemcpy(L.chunks, h.data(), sizeof(int64_t) * h.size(),
                      hipMemcpyHostToDevice));
}

void generate_natural(NaturalLevel &N, hipStream_t s) {
  generate_natural_part(N, N.nx, N.ny, N.nz, 0, 0, s);
}

void generate_natural_part(NaturalLevel &N, int nx, int ny, int nz, int64_t off,
                           int64_t offc, hipStream_t s) {
  const int64_t n = (int64_t)nx * ny * nz;
  HPCGX_REQUIRE(off >= 0 && off + n <= N.n, "grid rows outside the level");
  k_generate<<<grid_for(n), kBlock, 0, s>>>(nx, ny, nz, n, off, N.n, N.col, N.val,
                                            N.nnz_row);
  HPCGX_LAUNCHED();
  if (N.f2c) {
    const int64_t nc = n / 8;
    HPCGX_REQUIRE(offc >= 0 && offc + nc <= N.nc, "coarse rows outside the next level");
    k_f2c<<<grid_for(nc), kBlock, 0, s>>>(nx, ny, nc, nx / 2, ny / 2, off, offc, N.f2c);
    HPCGX_LAUNCHED();
  }
}

void finalize_level(Problem &P, int l, NaturalLevel &N, int flags,
                    hipStream_t s) {
  Level &L = P.L[l];
  const int64_t n = N.n;
  HPCGX_REQUIRE(n > 0 && n < ((int64_t)1 << 31), "rows per level must be in [1, 2^31)");
  L.nx = N.nx; L.ny = N.ny; L.nz = N.nz;
  L.n = n;

  // nnz
  {
    TempBuf cnt(sizeof(unsigned long long));
    HPCGX_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), s));
    k_count_nnz<<<grid_for(n), kBlock, 0, s>>>(n, N.nnz_row, cnt.as<unsigned long long>());
    HPCGX_LAUNCHED();
    unsigned long long v = 0;
    HPCGX_HIP(hipMemcpyAsync(&v, cnt.p, sizeof v, hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
    L.nnz = N.nnz = (int64_t)v;
  }

  // structure checks + lower-neighbour counts
  TempBuf remaining(sizeof(int32_t) * n), level(sizeof(int32_t) * n),
      color(sizeof(int32_t) * n), order(sizeof(int32_t) * n);
  {
    TempBuf bad(sizeof(int)), bw(sizeof(unsigned long long));
    HPCGX_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    HPCGX_HIP(hipMemsetAsync(bw.p, 0, sizeof(unsigned long long), s));
    k_check<<<grid_for(n), kBlock, 0, s>>>(n, N.col, N.nnz_row,
                                           remaining.as<int32_t>(), bad.as<int>(),
                                           bw.as<unsigned long long>());
    HPCGX_LAUNCHED();
    int b = 0;
    unsigned long long bwv = 0;
    HPCGX_HIP(hipMemcpyAsync(&b, bad.p, sizeof b, hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipMemcpyAsync(&bwv, bw.p, sizeof bwv, hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
    L.bandwidth = (int64_t)bwv;
    if (b & 1) throw Error(HPCGX_EINVAL, "column index out of range");
    if (b & 2) throw Error(HPCGX_EUNSUP, "a row has no diagonal entry");
    if (b & 4) throw Error(HPCGX_EUNSUP, "sparsity pattern is not symmetric");
  }

  // Levels and first-fit colours: a proposal from the rows' grids when it
  // passes the row-by-row check, else the Kahn traversal
  int nsched = 0;
  bool proposed = false;
  if (!N.parts.empty()) {
    const int np = (int)N.parts.size();
    TempBuf parts(sizeof(GridPart) * np), flags(2 * sizeof(int));
    HPCGX_HIP(hipMemcpyAsync(parts.p, N.parts.data(), sizeof(GridPart) * np,
                             hipMemcpyHostToDevice, s));
    HPCGX_HIP(hipMemsetAsync(flags.p, 0, 2 * sizeof(int), s));
    k_propose_schedule<<<grid_for(n), kBlock, 0, s>>>(n, np, parts.as<GridPart>(),
                                                      level.as<int32_t>(), color.as<int32_t>());
    HPCGX_LAUNCHED();
    k_verify_schedule<<<grid_for(n), kBlock, 0, s>>>(n, N.col, N.nnz_row, level.as<int32_t>(),
                                                     color.as<int32_t>(), flags.as<int>(),
                                                     flags.as<int>() + 1);
    HPCGX_LAUNCHED();
    int f[2] = {1, 0};
    HPCGX_HIP(hipMemcpyAsync(f, flags.p, sizeof f, hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
    proposed = f[0] == 0 && f[1] + 1 <= kMaxSched;
    if (proposed) nsched = f[1] + 1;
  }
  if (!proposed) {
    const int64_t cap = std::min<int64_t>(n, kMaxSched) + 2;
    TempBuf lptr(sizeof(int64_t) * cap), tail(sizeof(unsigned));
    HPCGX_HIP(hipMemsetAsync(lptr.p, 0, sizeof(int64_t) * cap, s));
    HPCGX_HIP(hipMemsetAsync(tail.p, 0, sizeof(unsigned), s));
    k_kahn_init<<<grid_for(n), kBlock, 0, s>>>(n, remaining.as<int32_t>(),
                                               order.as<int32_t>(), tail.as<unsigned>());
    HPCGX_LAUNCHED();
    // lptr[1] = rows of level 0
    k_kahn_advance<<<1, 1, 0, s>>>(-1, lptr.as<int64_t>(), tail.as<unsigned>());
    HPCGX_LAUNCHED();
    int t = 0;
    for (;;) {
      const int stop = (int)std::min<int64_t>(t + kSyncEvery, cap - 2);
      for (; t < stop; ++t) {
        k_kahn_step<<<kStepBlocks, kBlock, 0, s>>>(
            t, lptr.as<int64_t>(), n, N.col, remaining.as<int32_t>(),
            level.as<int32_t>(), color.as<int32_t>(), order.as<int32_t>(),
            tail.as<unsigned>());
        HPCGX_LAUNCHED();
        k_kahn_advance<<<1, 1, 0, s>>>(t, lptr.as<int64_t>(), tail.as<unsigned>());
        HPCGX_LAUNCHED();
      }
      // after step t-1, lptr[t+1] holds the end of level t
      int64_t last[2] = {0, 0};
      HPCGX_HIP(hipMemcpyAsync(last, lptr.as<int64_t>() + t, sizeof last,
                               hipMemcpyDeviceToHost, s));
      HPCGX_HIP(hipStreamSynchronize(s));
      if (last[1] == last[0]) break;  // level t is empty: traversal done
      if (t >= cap - 2) throw Error(HPCGX_EUNSUP, "too many schedule levels");
    }
    std::vector<int64_t> h(t + 2);
    HPCGX_HIP(hipMemcpy(h.data(), lptr.p, sizeof(int64_t) * (t + 2),
                        hipMemcpyDeviceToHost));
    int s_ = 0;
    while (s_ + 1 < (int)h.size() && h[s_ + 1] > h[s_]) ++s_;
    nsched = s_;
    if (h[nsched] != n)
      throw Error(HPCGX_EUNSUP, "dependency graph not traversable (cycle?)");
  }
  L.nsched = nsched;

  // device order = stable sort of rows by the colour's sweep position;
  // inside a colour, by band of grid lines when the level is an nx*ny*nz
  // grid in natural order
  L.iperm = P.arena.alloc<int32_t>(n);
  L.perm = P.arena.alloc<int32_t>(n);
  {
    int band = (flags >> 8) & 0xff;
    if (band == 0) band = kBandLines;
    const bool grid = (int64_t)N.nx * N.ny * N.nz == n && N.nx > 0;
    const int nbands = (grid && band > 1 && N.ny >= 2 * band) ? (N.ny + band - 1) / band : 1;
    // sweep position of each greedy colour
    TempBuf ckey(sizeof(int32_t) * n);
    {
      ColourRank rank;
      for (int c = 0; c < kMaxColors; ++c) rank.r[c] = c;
      const int sel = (flags >> 20) & 0xf;
      if (sel >= kNumSweepOrders) throw Error(HPCGX_EINVAL, "unknown sweep order");
      for (int q = 0; q < 8; ++q) rank.r[kSweepOrders[sel][q]] = q;
      k_colour_rank<<<grid_for(n), kBlock, 0, s>>>(n, color.as<int32_t>(), rank,
                                                   ckey.as<int32_t>());
      HPCGX_LAUNCHED();
    }
    std::vector<int64_t> cp;
    if (nbands > 1) {
      TempBuf key(sizeof(int32_t) * n);
      k_band_key<<<grid_for(n), kBlock, 0, s>>>(n, N.nx, N.ny, band, nbands,
                                                ckey.as<int32_t>(), key.as<int32_t>());
      HPCGX_LAUNCHED();
      const std::vector<int64_t> bp =
          sort_rows_by_key(n, key.as<int32_t>(), kMaxColors * nbands, L.iperm, s);
      for (int c = 0; c <= kMaxColors; ++c) cp.push_back(bp[(size_t)c * nbands]);
    } else {
      cp = sort_rows_by_key(n, ckey.as<int32_t>(), kMaxColors, L.iperm, s);
    }
    L.natural_in_colour = nbands == 1;
    std::vector<int64_t> used{cp[0]};  // the non-empty colours, in sweep order
    for (int c = 0; c < kMaxColors; ++c)
      if (cp[c + 1] > cp[c]) used.push_back(cp[c + 1]);
    L.color_ptr = used;
    L.ncolors = (int)used.size() - 1;
  }
  k_invert<<<grid_for(n), kBlock, 0, s>>>(n, L.iperm, L.perm);
  HPCGX_LAUNCHED();
  build_chunks(P, L, true);
  L.stream = (double)n * kSlots * (sizeof(double) + sizeof(int32_t)) > kStreamBytes;
  HPCGX_LAUNCHED();

  // lexicographic schedule: rows sorted by (level, row), renumbered
  L.level_rows = P.arena.alloc<int32_t>(n);
  L.level_ptr = sort_rows_by_key(n, level.as<int32_t>(), nsched, order.as<int32_t>(), s);
  k_map<<<grid_for(n), kBlock, 0, s>>>(n, L.perm, order.as<int32_t>(), L.level_rows);
  HPCGX_LAUNCHED();

  if (flags & HPCGX_KEEP_SCHEDULE) {
    L.level_of = P.arena.alloc<int32_t>(n);
    L.color_of = P.arena.alloc<int32_t>(n);
    HPCGX_HIP(hipMemcpyAsync(L.level_of, level.p, sizeof(int32_t) * n,
                             hipMemcpyDeviceToDevice, s));
    HPCGX_HIP(hipMemcpyAsync(L.color_of, color.p, sizeof(int32_t) * n,
                             hipMemcpyDeviceToDevice, s));
  }

  // ELL in device order
  L.col = P.arena.alloc<int32_t>((size_t)kSlots * n);
  L.val = P.arena.alloc<double>((size_t)kSlots * n);
  L.diag = P.arena.alloc<double>(n);
  L.dslot = P.arena.alloc<uint8_t>(n);
  L.nnz_row = P.arena.alloc<uint8_t>(n);
  k_permute_ell<<<grid_for(n), kBlock, 0, s>>>(n, L.iperm, L.perm, N.col, N.val,
                                               N.nnz_row, L.col, L.val, L.diag,
                                               L.dslot, L.nnz_row);
  HPCGX_LAUNCHED();
  // 16-bit column ids where every colour block starts on a tile boundary
  // (the colour-step and SpMV launches then keep each wavefront in a tile)
  bool aligned = n >= kTile;
  for (int c = 0; c <= L.ncolors; ++c) aligned = aligned && L.color_ptr[c] % kTile == 0;
  if (aligned) {
    const int64_t T = (n + kTile - 1) / kTile;
    L.ntiles = T;
    L.col16 = P.arena.alloc<uint16_t>((size_t)kSlots * n);
    L.cbase = P.arena.alloc<int32_t>((size_t)kSlots * T);
    L.tbase = P.arena.alloc<int32_t>((size_t)kBaseStride * T);
    L.tinfo = P.arena.alloc<int32_t>(T);
    L.amask = P.arena.alloc<uint8_t>(T);
    L.off8 = P.arena.alloc<uint32_t>((size_t)kSlotWords * n);
    const int64_t wpb = kBlock / 64;
    k_tile_layout<<<(unsigned)((T + wpb - 1) / wpb), kBlock, 0, s>>>(
        n, T, L.iperm, L.col, L.val, L.nnz_row, L.dslot, L.col16, L.off8, L.cbase, L.tbase, L.tinfo,
        L.amask);
    HPCGX_LAUNCHED();
    std::vector<int32_t> info(T);
    HPCGX_HIP(hipMemcpyAsync(info.data(), L.tinfo, sizeof(int32_t) * T,
                             hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
    L.wide_tiles = std::count(info.begin(), info.end(), (int32_t)kTile32);
    L.byte_tiles = std::count(info.begin(), info.end(), (int32_t)kTile8);
    L.c16 = true;
    L.zmask = P.arena.alloc<uint32_t>(T);
    TempBuf cp(sizeof(int64_t) * (L.ncolors + 1));
    HPCGX_HIP(hipMemcpyAsync(cp.p, L.color_ptr.data(), sizeof(int64_t) * (L.ncolors + 1),
                             hipMemcpyHostToDevice, s));
    k_zero_mask<<<(unsigned)((T + wpb - 1) / wpb), kBlock, 0, s>>>(
        n, T, L.ncolors, cp.as<int64_t>(), L.col, L.val, L.zmask);
    HPCGX_LAUNCHED();
    HPCGX_HIP(hipStreamSynchronize(s));
  }
  HPCGX_HIP(hipStreamSynchronize(s));
  if (P.use_wave && n >= kWaveMinRows)
    build_wave(P, L, kWaveSlabFactor, kWaveRows, s);
}

// Wavefront schedule of the multicolour sweeps.  Natural rows are cut into
// slabs of W >= bandwidth rows, so a row of slab s only couples to slabs
// s-1..s+1.  Tile (c, s) = device rows of colour c in slab s: contiguous,
// because a colour block keeps natural order.  Tasks (chunks of kBlock rows
// of one tile) are listed by t = s + 2*phase, then phase: a task of phase p
// at slab s needs only the phase p-1 tiles at slabs s-1..s+1 (everything it
// must see updated, or must read before it is updated, follows from those),
// and each of them comes earlier in the list.
void drop_wave(Problem &P, Level &L) {
  const int64_t C = L.ncolors, S = L.nslabs;
  if (L.tile_ptr) P.arena.release(L.tile_ptr, sizeof(int64_t) * C * (S + 1));
  for (int d = 0; d < 2; ++d)
    if (L.tasks[d]) P.arena.release(L.tasks[d], sizeof(int32_t) * 2 * L.ntasks[d]);
  if (L.done) P.arena.release(L.done, sizeof(unsigned) * C * S);
  L.tile_ptr = nullptr;
  L.tasks[0] = L.tasks[1] = nullptr;
  L.ntasks[0] = L.ntasks[1] = 0;
  L.done = nullptr;
  L.nslabs = 0;
  L.wave = false;
}

void build_wave(Problem &P, Level &L, int slab_factor, int rows, hipStream_t s) {
  if (!L.natural_in_colour)
    throw Error(HPCGX_EUNSUP,
                "wavefront sweeps need natural row order inside each colour "
                "(create the problem with HPCGX_BAND_LINES(1))");
  drop_wave(P, L);
  const int C = L.ncolors;
  const int CH = kBlock * std::max(1, rows);
  L.wave_chunk = CH;
  const int64_t n = L.n;
  const int64_t W = std::max<int64_t>(1, L.bandwidth) * std::max(1, slab_factor);
  const int S = (int)((n + W - 1) / W);
  L.wave_slab = W;
  L.nslabs = S;
  const int64_t ntp = (int64_t)C * (S + 1);
  L.tile_ptr = P.arena.alloc<int64_t>(ntp);
  TempBuf cp(sizeof(int64_t) * (C + 1));
  HPCGX_HIP(hipMemcpyAsync(cp.p, L.color_ptr.data(), sizeof(int64_t) * (C + 1),
                           hipMemcpyHostToDevice, s));
  k_tiles<<<grid_for(ntp), kBlock, 0, s>>>(C, S, W, cp.as<int64_t>(), L.iperm,
                                           L.tile_ptr);
  HPCGX_LAUNCHED();
  std::vector<int64_t> tp(ntp);
  HPCGX_HIP(hipMemcpyAsync(tp.data(), L.tile_ptr, sizeof(int64_t) * ntp,
                           hipMemcpyDeviceToHost, s));
  HPCGX_HIP(hipStreamSynchronize(s));
  for (int dir = 0; dir < 2; ++dir) {
    std::vector<int32_t> tasks;
    for (int t = 0; t < S + 2 * (C - 1); ++t)
      for (int p = 0; p < C; ++p) {
        const int sl = t - 2 * p;
        if (sl < 0 || sl >= S) continue;
        const int c = dir == 0 ? p : C - 1 - p;
        const int64_t nrows = tp[(int64_t)c * (S + 1) + sl + 1] - tp[(int64_t)c * (S + 1) + sl];
        const int64_t nch = (nrows + CH - 1) / CH;
        for (int64_t k = 0; k < nch; ++k) {
          tasks.push_back(p * S + sl);
          tasks.push_back((int32_t)k);
        }
      }
    L.ntasks[dir] = (int64_t)tasks.size() / 2;
    L.tasks[dir] = P.arena.alloc<int32_t>(tasks.size());
    HPCGX_HIP(hipMemcpy(L.tasks[dir], tasks.data(), sizeof(int32_t) * tasks.size(),
                        hipMemcpyHostToDevice));
  }
  L.done = P.arena.alloc<unsigned>((size_t)C * S);
  HPCGX_HIP(hipMemset(L.done, 0, sizeof(unsigned) * C * S));
  L.wave = true;
}

void link_levels(Problem &P, int l, const NaturalLevel &Nf, hipStream_t s) {
  Level &F = P.L[l];
  const Level &C = P.L[l + 1];
  F.nc = C.n;
  F.f2c = P.arena.alloc<int32_t>(C.n);
  F.rc = P.arena.alloc<double>(C.n);
  F.xc = P.arena.alloc<double>(C.n);
  k_link<<<grid_for(C.n), kBlock, 0, s>>>(C.n, C.iperm, Nf.f2c, F.perm, F.f2c);
  HPCGX_LAUNCHED();
  // Fine-order transfers (Level::ftile, rperm): the fine rows f2c reaches,
  // sorted, with their coarse rows; kept when every kTile of them (the last
  // group possibly fewer) are consecutive rows from a multiple of kTile.
  const int64_t nc = C.n;
  std::vector<int32_t> fr(nc), cr(nc);
  {
    TempBuf ids(sizeof(int32_t) * nc), fs(sizeof(int32_t) * nc), cs(sizeof(int32_t) * nc);
    k_iota<<<grid_for(nc), kBlock, 0, s>>>(nc, ids.as<int32_t>());
    HPCGX_LAUNCHED();
    const uint32_t *key = reinterpret_cast<const uint32_t *>(F.f2c);
    size_t bytes = 0;
    HPCGX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key, fs.as<uint32_t>(),
                                                 ids.as<int32_t>(), cs.as<int32_t>(), (int)nc, 0,
                                                 bits_for(F.n), s));
    TempBuf temp(bytes);
    HPCGX_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, bytes, key, fs.as<uint32_t>(),
                                                 ids.as<int32_t>(), cs.as<int32_t>(), (int)nc, 0,
                                                 bits_for(F.n), s));
    HPCGX_HIP(hipMemcpyAsync(fr.data(), fs.p, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipMemcpyAsync(cr.data(), cs.p, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
  }
  const int64_t ng = (nc + kTile - 1) / kTile;
  std::vector<int32_t> ft(ng);
  bool tiled = true;
  for (int64_t g = 0; g < ng && tiled; ++g) {
    const int64_t f0 = fr[g * kTile];
    tiled = f0 % kTile == 0 && (g == 0 || f0 > fr[g * kTile - 1]);
    for (int64_t t = g * kTile + 1; t < std::min(nc, (g + 1) * kTile) && tiled; ++t)
      tiled = fr[t] == f0 + (t - g * kTile);
    ft[g] = (int32_t)(f0 / kTile);
  }
  if (tiled) {
    F.ftile = P.arena.alloc<int32_t>(ng);
    F.rperm = P.arena.alloc<int32_t>(nc);
    HPCGX_HIP(hipMemcpy(F.ftile, ft.data(), sizeof(int32_t) * ng, hipMemcpyHostToDevice));
    HPCGX_HIP(hipMemcpy(F.rperm, cr.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice));
  }
}

namespace {
// CheckProblem.cpp:79-148 on the device, for a level in device order whose
// rows come from known grids.  Device row i is natural row g = iperm[i] of
// part p; walking its stencil neighbours in the generator's order (sz, sy,
// sx ascending, so ascending columns), the k-th neighbour must be the
// row's k-th stored entry (slots in order, skipping the exact zeros of the
// aligned ELL's padding): natural column iperm[col] == the neighbour, value
// 26 on the diagonal and -1 elsewhere; nothing stored after the last one;
// nonzerosInRow == the neighbour count; the diagonal array holds 26.  With
// f2c (every level but the coarsest), coarse row c of part p injects from
// the fine point (2x, 2y, 2z) (GenerateCoarseProblem.cpp:71-82).
// Failures set bits of *flags (1 value, 2 column, 4 count, 8 extra entry,
// 16 diagonal, 32 f2c, 64 row outside every grid) and the smallest failing
// natural row in *first; *nnz sums the neighbour counts.
struct CheckOut {
  int flags;
  int pad;
  unsigned long long first;
  unsigned long long nnz;
  unsigned long long bad_rows;
};

__device__ inline int part_of(int nparts, const GridPart *__restrict__ parts, int64_t g) {
  int lo = 0, hi = nparts - 1;  // last part with off <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (parts[mid].off <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void k_check_problem(
    int64_t n, int nparts, const GridPart *__restrict__ parts, const int32_t *__restrict__ iperm,
    const int32_t *__restrict__ perm, const int32_t *__restrict__ col,
    const double *__restrict__ val, const uint8_t *__restrict__ nnz_row,
    const double *__restrict__ diag, CheckOut *out) {
  const int64_t i = gtid();
  int flags = 0;
  unsigned long long cnt = 0, g = ~0ull;
  if (i < n) {
    g = (unsigned long long)iperm[i];
    const GridPart p = parts[part_of(nparts, parts, (int64_t)g)];
    const int64_t local = (int64_t)g - p.off;
    if ((int64_t)g >= n || perm[g] != i || local < 0 ||
        local >= (int64_t)p.nx * p.ny * p.nz) {
      flags = 64;
    } else {
      const int ix = (int)(local % p.nx);
      const int64_t t = local / p.nx;
      const int iy = (int)(t % p.ny), iz = (int)(t / p.ny);
      const int64_t nxy = (int64_t)p.nx * p.ny;
      int j = 0;
      for (int sz = -1; sz <= 1; ++sz) {
        if (iz + sz < 0 || iz + sz >= p.nz) continue;
        for (int sy = -1; sy <= 1; ++sy) {
          if (iy + sy < 0 || iy + sy >= p.ny) continue;
          for (int sx = -1; sx <= 1; ++sx) {
            if (ix + sx < 0 || ix + sx >= p.nx) continue;
            const int64_t want = (int64_t)g + sz * nxy + (int64_t)sy * p.nx + sx;
            while (j < kSlots && val[(int64_t)j * n + i] == 0.0) ++j;
            ++cnt;
            if (j == kSlots) { flags |= 4; continue; }
            const int64_t k = (int64_t)j * n + i;
            const int32_t c = col[k];
            if (c < 0 || c >= n || iperm[c] != want) flags |= 2;
            if (val[k] != (want == (int64_t)g ? 26.0 : -1.0)) flags |= 1;
            ++j;
          }
        }
      }
      for (; j < kSlots; ++j)
        if (val[(int64_t)j * n + i] != 0.0) flags |= 8;
      if (nnz_row[i] != cnt) flags |= 4;
      if (diag[i] != 26.0) flags |= 16;
    }
  }
  if (flags) {
    atomicOr(&out->flags, flags);
    atomicMin(&out->first, g);
    atomicAdd(&out->bad_rows, 1ull);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(&out->nnz, cnt);
}

// f2c of GenerateCoarseProblem.cpp:77-79 for coarse device rows: natural
// fine row iperm_f[f2c[c]] must be the fine point (2x, 2y, 2z) of natural
// coarse row iperm_c[c], in the same part (fine part q at foff[q]).
__global__ __launch_bounds__(kBlock) void k_check_f2c(
    int64_t nc, int nparts, const GridPart *__restrict__ cparts,
    const GridPart *__restrict__ fparts, const int32_t *__restrict__ iperm_c,
    const int32_t *__restrict__ iperm_f, const int32_t *__restrict__ f2c, int64_t nf,
    CheckOut *out) {
  const int64_t c = gtid();
  if (c >= nc) return;
  const int64_t gc = iperm_c[c];
  const int q = part_of(nparts, cparts, gc);
  const GridPart pc = cparts[q], pf = fparts[q];
  const int64_t local = gc - pc.off;
  bool ok = local >= 0 && local < (int64_t)pc.nx * pc.ny * pc.nz;
  if (ok) {
    const int ix = (int)(local % pc.nx);
    const int64_t t = local / pc.nx;
    const int iy = (int)(t % pc.ny), iz = (int)(t / pc.ny);
    const int64_t want = pf.off + ((int64_t)(2 * iz) * pf.ny + 2 * iy) * pf.nx + 2 * ix;
    const int32_t f = f2c[c];
    ok = f >= 0 && f < nf && iperm_f[f] == want;
  }
  if (!ok) {
    atomicOr(&out->flags, 32);
    atomicMin(&out->first, (unsigned long long)gc);
    atomicAdd(&out->bad_rows, 1ull);
  }
}

// b, x, xexact of GenerateProblem_ref.cpp:178-180 in device order on the
// fine level: b_i = 26 - (nnz_i - 1) with nnz_i the stencil's neighbour
// count, x_i = 0, xexact_i = 1 (CheckProblem.cpp:139-141).
__device__ inline int axis_count(int v, int n) {
  return 1 + (v > 0) + (v + 1 < n);
}

__global__ __launch_bounds__(kBlock) void k_check_vectors(
    int64_t n, int nparts, const GridPart *__restrict__ parts, const int32_t *__restrict__ iperm,
    const double *b, const double *x, const double *xexact, CheckOut *out) {
  const int64_t i = gtid();
  if (i >= n) return;
  const int64_t g = iperm[i];
  const GridPart p = parts[part_of(nparts, parts, g)];
  const int64_t local = g - p.off;
  const int ix = (int)(local % p.nx);
  const int64_t t = local / p.nx;
  const int iy = (int)(t % p.ny), iz = (int)(t / p.ny);
  const int m = axis_count(ix, p.nx) * axis_count(iy, p.ny) * axis_count(iz, p.nz);
  int flags = 0;
  if (b && b[i] != 26.0 - (double)(m - 1)) flags |= 128;
  if (x && x[i] != 0.0) flags |= 256;
  if (xexact && xexact[i] != 1.0) flags |= 512;
  if (flags) {
    atomicOr(&out->flags, flags);
    atomicMin(&out->first, (unsigned long long)g);
    atomicAdd(&out->bad_rows, 1ull);
  }
}

std::string check_message(int l, const CheckOut &o) {
  static const char *what[] = {"a value differs",       "a column differs",
                               "an entry count differs", "an entry beyond the stencil",
                               "a diagonal differs",    "f2c differs",
                               "a row outside the grids", "b differs",
                               "x is not 0",            "xexact is not 1"};
  std::string m = "CheckProblem: level " + std::to_string(l) + ": " +
                  std::to_string(o.bad_rows) + " rows differ from the generated problem (first: "
                  "natural row " + std::to_string(o.first) + ";";
  for (int b = 0; b < 10; ++b)
    if (o.flags & (1 << b)) m += std::string(" ") + what[b] + ";";
  m.back() = ')';
  return m;
}
}  // namespace

// The grids level l's rows come from: one per problem (fleets: problem q at
// its natural row offset), or none when a level's grid is unknown.
std::vector<GridPart> grid_parts(const Problem &P, int l) {
  std::vector<GridPart> parts;
  const Level &L = P.L[l];
  if (P.fleet.count > 0) {
    for (int q = 0; q < P.fleet.count; ++q)
      parts.push_back(GridPart{P.fleet.off[l][q], P.fleet.dims[3 * q] >> l,
                               P.fleet.dims[3 * q + 1] >> l, P.fleet.dims[3 * q + 2] >> l});
  } else if ((int64_t)L.nx * L.ny * L.nz == L.n && L.n > 0) {
    parts.push_back(GridPart{0, L.nx, L.ny, L.nz});
  }
  return parts;
}

void check_problem(const Problem &P, const double *b, const double *x, const double *xexact,
                   hipStream_t s) {
  std::vector<std::vector<GridPart>> parts(P.nlevels);
  for (int l = 0; l < P.nlevels; ++l) {
    parts[l] = grid_parts(P, l);
    if (parts[l].empty())
      throw Error(HPCGX_EINVAL, "CheckProblem: the grid of level " + std::to_string(l) +
                                    " is unknown (its rows do not come from nx*ny*nz)");
  }
  CheckOut init{0, 0, ~0ull, 0ull, 0ull};
  for (int l = 0; l < P.nlevels; ++l) {
    const Level &L = P.L[l];
    const int np = (int)parts[l].size();
    TempBuf dp(sizeof(GridPart) * np), dpc(sizeof(GridPart) * np), dout(sizeof(CheckOut));
    HPCGX_HIP(hipMemcpyAsync(dp.p, parts[l].data(), sizeof(GridPart) * np, hipMemcpyHostToDevice,
                             s));
    HPCGX_HIP(hipMemcpyAsync(dout.p, &init, sizeof init, hipMemcpyHostToDevice, s));
    k_check_problem<<<grid_for(L.n), kBlock, 0, s>>>(L.n, np, dp.as<GridPart>(), L.iperm, L.perm,
                                                     L.col, L.val, L.nnz_row, L.diag,
                                                     dout.as<CheckOut>());
    HPCGX_LAUNCHED();
    if (l + 1 < P.nlevels) {
      const Level &C = P.L[l + 1];
      HPCGX_REQUIRE(L.f2c && parts[l + 1].size() == parts[l].size(),
                    "CheckProblem: the levels' grids do not correspond");
      HPCGX_HIP(hipMemcpyAsync(dpc.p, parts[l + 1].data(), sizeof(GridPart) * np,
                               hipMemcpyHostToDevice, s));
      k_check_f2c<<<grid_for(C.n), kBlock, 0, s>>>(C.n, np, dpc.as<GridPart>(), dp.as<GridPart>(),
                                                   C.iperm, L.iperm, L.f2c, L.n,
                                                   dout.as<CheckOut>());
      HPCGX_LAUNCHED();
    }
    if (l == 0 && (b || x || xexact)) {
      k_check_vectors<<<grid_for(L.n), kBlock, 0, s>>>(L.n, np, dp.as<GridPart>(), L.iperm, b, x,
                                                       xexact, dout.as<CheckOut>());
      HPCGX_LAUNCHED();
    }
    CheckOut o;
    HPCGX_HIP(hipMemcpyAsync(&o, dout.p, sizeof o, hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
    int64_t want_nnz = 0;  // CheckProblem.cpp:145-148: prod(3 n_a - 2) per grid
    for (const GridPart &g : parts[l])
      want_nnz += (int64_t)(3 * g.nx - 2) * (3 * g.ny - 2) * (3 * g.nz - 2);
    if (o.flags) throw Error(HPCGX_ECHECK, check_message(l, o));
    if ((int64_t)o.nnz != want_nnz || L.nnz != want_nnz)
      throw Error(HPCGX_ECHECK, "CheckProblem: level " + std::to_string(l) + ": " +
                                    std::to_string(L.nnz) + " nonzeros, the grids give " +
                                    std::to_string(want_nnz));
  }
}

void problem_vectors(const Problem &P, double *b, double *x, double *xexact,
                     hipStream_t s) {
  const Level &L = P.L[0];
  k_vectors<<<grid_for(L.n), kBlock, 0, s>>>(L.n, L.nnz_row, b, x, xexact);
  HPCGX_LAUNCHED();
}

}  // namespace hpcgx
