// internal.hpp -- device-side state of libhpcg_amd and the kernel launchers
// shared between translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "hpcg_amd.h"

namespace hpcgx {

constexpr int kSlots = 27;        // ELL width: the 27-point stencil's maximum
constexpr int kMaxLevels = 8;
constexpr int kMaxColors = 32;    // first-fit colours of a <=27-entry row: <=27
constexpr int kBlock = 256;       // threads per workgroup for streaming kernels
// A launch holds fewer than 2^32 work-items (the dispatch packet's grid
// size is 32 bits): kernels over all n*27 ELL entries loop over the slots
// inside one work-item per row.
constexpr int kReduceBlocks = 4096;  // workgroups of a grid-stride reduction
constexpr int kMaxPartials = 8192;   // partial-sum slots (largest reducing grid)
constexpr double kStreamBytes = 512e6;  // matrices above this are streamed
constexpr int64_t kWaveMinRows = 1 << 23;  // levels swept as a wavefront
constexpr int kWaveSlabFactor = 2;         // slab width / matrix bandwidth
constexpr int kWaveRows = 2;               // rows per thread of a wavefront task
constexpr int kTile = 64;  // rows sharing narrow column bases (one wavefront)
constexpr int kSlotWords = (kSlots + 3) / 4;  // 8-bit offsets of 4 slots per word
constexpr int kPairWords = (kSlots + 1) / 2;  // FP16 values of 2 slots per word
constexpr int kBaseStride = 32;  // ints per tile in Level::tbase (128 bytes)
enum TileClass { kTile8 = 0, kTile32 = 1, kTile16 = 2 };
constexpr int kBandLines = 1;  // grid lines per band of the in-colour order (1 = natural)

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HPCGX_HIP(call)                                                       \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::hpcgx::Error(e_ == hipErrorOutOfMemory ? HPCGX_ENOMEM           \
                                                     : HPCGX_EHIP,            \
                           std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define HPCGX_LAUNCHED() HPCGX_HIP(hipGetLastError())

// The calling thread's message for hpcgx_last_error (capi.hip).
std::string &last_error_slot();

// Runs f for an extern "C" entry point: exceptions never cross the
// boundary; each maps to an HPCGX_E* code and its message is recorded.
template <class F>
int guard(F &&f) {
  try {
    f();
    return HPCGX_OK;
  } catch (const Error &e) {
    last_error_slot() = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    last_error_slot() = "host allocation failed";
    return HPCGX_ENOMEM;
  } catch (const std::exception &e) {
    last_error_slot() = e.what();
    return HPCGX_EHIP;
  } catch (...) {
    last_error_slot() = "unknown failure";
    return HPCGX_EHIP;
  }
}

#define HPCGX_REQUIRE(cond, msg)                                  \
  do {                                                            \
    if (!(cond)) throw ::hpcgx::Error(HPCGX_EINVAL, (msg));       \
  } while (0)

// Device allocation tracked for OptimizeProblemMemoryUse.
struct DeviceArena {
  std::vector<void *> ptrs;
  double bytes = 0.0;
  template <class T>
  T *alloc(size_t count) {
    void *p = nullptr;
    if (count == 0) count = 1;
    HPCGX_HIP(hipMalloc(&p, count * sizeof(T)));
    ptrs.push_back(p);
    bytes += double(count) * sizeof(T);
    return static_cast<T *>(p);
  }
  void release(void *p, size_t bytes_of) {
    for (auto &q : ptrs)
      if (q == p) {
        (void)hipFree(q);
        q = nullptr;
        bytes -= double(bytes_of);
        return;
      }
  }
  ~DeviceArena() {
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
  }
};

// One level of the hierarchy.  Rows live in DEVICE ORDER: sorted by colour,
// natural order within a colour.  ELL storage is column-major: slot j of
// device row i is at [j * n + i]; unused slots hold (col = i, val = 0) and
// sit after the row's stored entries, so a full 27-slot traversal adds exact
// zeros after the reference's terms.
struct Level {
  bool finest = false;  // level 0 (its kernels get their own symbols)
  bool stream = false;  // matrix >> last-level cache: non-temporal matrix loads
  bool xcd = false;     // XCD-aware workgroup -> rows mapping (measured slower)
  bool pair_rows = false;  // colour steps with two adjacent rows per thread
  int nx = 0, ny = 0, nz = 0;
  int64_t n = 0, nnz = 0;
  int32_t *col = nullptr;   // [kSlots * n] device-order column ids
  double *val = nullptr;    // [kSlots * n]
  double *diag = nullptr;   // [n]
  uint8_t *dslot = nullptr; // [n] slot of the diagonal entry
  uint8_t *nnz_row = nullptr; // [n] stored entries per row
  int32_t *perm = nullptr;  // natural row -> device row
  int32_t *iperm = nullptr; // device row -> natural row
  // multicolour schedule: colour c owns device rows [color_ptr[c], color_ptr[c+1])
  // in natural order (natural_in_colour) or band by band (k_band_key)
  bool natural_in_colour = true;
  int ncolors = 0;
  std::vector<int64_t> color_ptr;
  // lexicographic schedule: level s owns level_rows[level_ptr[s] .. level_ptr[s+1])
  int nsched = 0;
  std::vector<int64_t> level_ptr;
  int32_t *level_rows = nullptr;  // device-order row ids, level-major
  // row chunks (<= kBlock rows, [begin, end) pairs) interleaving the colours
  int64_t nchunks = 0;
  int64_t *chunks = nullptr;
  bool chunks_interleaved = true;  // else colour block by colour block
  // wavefront multicolour sweeps: tile (c, s) = device rows of colour c in
  // natural-row slab s = [tile_ptr[c*(nslabs+1)+s], tile_ptr[c*(nslabs+1)+s+1])
  int64_t bandwidth = 0;  // max |col - row| in natural order
  bool wave = false;
  int64_t wave_slab = 0;
  int wave_chunk = kBlock;  // rows per task
  int nslabs = 0;
  int64_t *tile_ptr = nullptr;
  int64_t ntasks[2] = {0, 0};       // forward, backward
  int32_t *tasks[2] = {nullptr, nullptr};  // (phase*nslabs + slab, chunk) pairs
  unsigned *done = nullptr;          // [ncolors*nslabs] finished chunks per tile
  // optional per-row ids (natural order) kept for inspection
  int32_t *level_of = nullptr, *color_of = nullptr;
  // mixed precision (built on first use): FP16 values with the diagonal slot
  // zeroed, the FP32 diagonal, FP32 coarse-level vectors
  _Float16 *val16 = nullptr;
  uint32_t *vc = nullptr;  // c16 levels: col16 << 16 | FP16 bits, one word per slot
  uint32_t *v16p = nullptr;  // c16 levels: FP16 bits of slots 2w, 2w+1 in word w ([kPairWords][n])
  float *diag32 = nullptr;
  float *rc32 = nullptr, *xc32 = nullptr;
  // Narrow column ids.  Tile t = device rows [kTile t, kTile t + kTile).  The
  // slots of a tile are aligned: slot j holds the tile's j-th distinct
  // natural column offset (col - row in natural numbering, ascending); a row
  // without that offset holds a valid column and the value 0 there, so each
  // row still meets its stored entries in ascending column order and the
  // extra terms add exact zeros.  Column of slot j of row i in tile t:
  //   tinfo[t] == kTile8:  cbase[j * ntiles + t] + byte (j % 4) of
  //                        off8[(j / 4) * n + i]   (every slot spans <= 255)
  //   tinfo[t] == kTile16: cbase[j * ntiles + t] + col16[j * n + i]
  // tbase[t * kBaseStride + j] holds the same bases tile by tile.  The FP64
  // kernels read cbase, whose lines serve 16 neighbouring tiles; the FP32
  // kernels read tbase, where a tile's bases come in a few wide scalar loads
  // (measured: each is the faster layout for its own kernels).
  //   tinfo[t] == kTile32: col[j * n + i] (more than 27 distinct offsets, or
  //                        a slot spanning more than 65535 columns)
  // `col` holds the 32-bit ids of every tile, col16 those of every 8- and
  // 16-bit tile.
  bool c16 = false;
  int64_t ntiles = 0;
  int64_t wide_tiles = 0;   // kTile32 tiles
  int64_t byte_tiles = 0;   // kTile8 tiles
  uint16_t *col16 = nullptr;
  uint32_t *off8 = nullptr;  // [kSlotWords * n]
  // per kTile8 tile: bit g set when every row's offset word g is
  // lane * 0x01010101 (each of its slots' columns is the slot base plus the
  // row's position in the tile), so the kernels need not load it
  uint8_t *amask = nullptr;
  int32_t *cbase = nullptr;
  int32_t *tbase = nullptr;
  int32_t *tinfo = nullptr;
  // Per tile, bit j set when some row of the tile stores a nonzero in slot j
  // whose column lies in a colour visited before the row's own (a device
  // row below the first row of the row's colour): the only slots the first
  // forward sweep from x = 0 reads (launch_symgs_rows_zero).
  uint32_t *zmask = nullptr;
  // transfer to the next level: f2c[coarse device row] = fine device row
  int64_t nc = 0;
  int32_t *f2c = nullptr;
  // When the fine rows f2c reaches, in ascending order, form runs of
  // consecutive rows that each start at a multiple of kTile and (all but the
  // last) hold a multiple of kTile rows (the injected points of the 27-point
  // grid are one parity class, a colour block), the transfers run in fine
  // order: thread t owns fine row fine_row(ftile, t) (ascending in t) and
  // coarse row rperm[t], so a wavefront reads whole tiles of matrix rows.
  int32_t *ftile = nullptr;  // per kTile transfer rows: fine row / kTile
  int32_t *rperm = nullptr;
  double *rc = nullptr, *xc = nullptr;  // next level's residual and correction
};

struct Problem {
  int device = 0;
  bool generated = false;  // built by hpcgx_generate_problem
  int nlevels = 0;
  Level L[kMaxLevels];
  DeviceArena arena;
  // CG work vectors (CGData.hpp:27-32 of the reference)
  double *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr;
  // reduction scratch: partial sums, completion ticket, scalars
  double *partials = nullptr;
  unsigned *ticket = nullptr;
  double *scalars = nullptr;  // see CgScalar
  int *wave_err = nullptr;    // set when a wavefront wait gives up
  int wave_grid = 0;          // co-resident workgroups of the wavefront kernel
  bool use_wave = false;      // multicolour sweeps as wavefronts (measured slower)
  bool use_c16 = true;        // read narrow column ids where a level has them
  bool use_c8 = true;         // ... 8-bit ones where a tile has them
  bool use_pairs = true;      // FP32 colour steps of 8-bit tiles read FP16 value pairs
  bool use_affine = true;     // skip the offset words Level::amask marks ("affine8")
  bool stream_p = false;      // p update with non-temporal accesses ("stream_p")
  bool zero_guess = true;     // V-cycle's first forward sweeps skip the zero x ("zero_guess")
  bool defer_x = true;        // CG takes x's step with the next p update ("defer_x")
  int gs_block = kBlock;      // workgroup size of the finest colour steps ("gs_block")
  int gs_waves = 1;           // 1: register caps on the finest colour steps ("gs_waves")
  // finest colour steps and SpMV: on affine tiles, slot pairs whose x windows
  // are one element apart share one gather ("x_pairs"; bit 0: the colour
  // steps without r.z and the SpMV, bit 1: the r.z colour step)
  int x_pairs = 3;
  // the r (and x) update of CG with one row per thread and a second launch
  // for r.r ("xr_rows")
  bool xr_rows = true;
  // zero-x colour steps issue loads for the live slots only ("zero_live"):
  // bit 0 the FP64 finest level, bit 1 the FP64 coarser levels, bit 2 the
  // FP32 V-cycle, bit 3 the block V-cycle
  int zero_live = 11;
  // the mixed CG's FP64 residual recompute in one pass ("resid_fused")
  bool resid_fused = true;
  // the mixed CG's FP64 A p reads 8-bit tiles as FP16 value pairs and 8-bit
  // offsets, not 4-byte offset/value words ("hc_pairs")
  bool hc_pairs = true;
  // non-temporal multivector loads and stores in the block CG's vector
  // kernels ("block_nt")
  bool block_nt = true;
  // block PCG finest colour steps: slot 3m + 2's x rows from slot 3m's by
  // W-lane moves on affine tiles ("block_pairs")
  bool block_pairs = false;
  int spmv_dot_blocks = 4096; // persistent grid of the fused SpMV + dot
  // The fused SpMV + dot with one workgroup per row chunk of the fine level
  // ("spmv_dot_chunks"): the p.Ap partial of each chunk, summed in fixed
  // order by a second launch.
  bool spmv_dot_chunked = true;
  int64_t nchunk_partials = 0;
  double *chunk_partials = nullptr;   // [nchunk_partials]
  // r.z from partials that the finest level's last colour sweep leaves
  // ("rz_fused"), summed in fixed order by a second launch.
  bool rz_fused = true;
  int64_t nrz_partials = 0;
  double *rz_part = nullptr;          // [nrz_partials]
  double *dot_parts = nullptr;        // [kDotParts]: launch_dot's first pass
  bool mixed_ready = false;   // FP16/FP32 copies built
  bool half_exact = false;    // every off-diagonal value is exact in FP16
  float *r32 = nullptr, *z32 = nullptr;  // FP32 input / output of the V-cycle
  // block PCG over s right-hand sides (block.hip), allocated on first use:
  // [n][s] multivectors of the fine level, FP32 coarse multivectors,
  // per-right-hand-side scalars (rtz, old rtz, alpha, beta, |r|^2: s each)
  struct BlockState {
    int s = 0;
    double *R = nullptr, *Pv = nullptr, *AP = nullptr;
    float *R32 = nullptr, *Z32 = nullptr;
    float *rc32[kMaxLevels] = {}, *xc32[kMaxLevels] = {};
    double *partials = nullptr;
    double *cpart = nullptr;  // per-workgroup column sums of k_bspmv_chunks
    int64_t ncpart = 0;
    unsigned *ticket = nullptr;
    double *scalars = nullptr;
    void release(Problem &P);
  } blk;
  // captured CG iteration (HPCGX_CG_GRAPH): body of iterations 2.. for one
  // (x, ordering, preconditioning) key; captured on a private stream
  struct IterGraph {
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
    const double *x = nullptr;
    int ordering = -1;
    bool precond = false;
    void reset() {
      if (exec) (void)hipGraphExecDestroy(exec);
      exec = nullptr;
      x = nullptr;
    }
    ~IterGraph() {
      reset();
      if (cap) (void)hipStreamDestroy(cap);
    }
  } graph;
  // captured body of the mixed-precision CG's iterations 2.. (mixed.hip),
  // keyed by x and the ordering; used while use_mixed_graph is set
  IterGraph mgraph;
  bool use_mixed_graph = true;
  // captured body of the block CG's iterations 2.. (block.hip), keyed by X,
  // the ordering and s; dropped when the block state is reallocated
  IterGraph bgraph;
  int bgraph_s = 0;
  // A fleet: the block-diagonal union of `count` independent problems
  // (fleet.hip).  Level l holds every problem's level l, natural rows
  // concatenated problem by problem (problem q: rows [off[l][q], off[l][q+1])).
  // Setup, SpMV and the V-cycle treat it as one matrix; CG keeps its scalars
  // per problem.  The fine level's device rows are cut into pieces of at most
  // kPieceRows rows inside one (problem, colour) range, listed problem by
  // problem: pieces[3k..3k+2] = begin, end, problem; problem q owns pieces
  // [pptr[q], pptr[q+1]).
  struct FleetState {
    int count = 0;  // 0: not a fleet
    std::vector<int> dims;  // nx, ny, nz of each problem
    std::vector<int64_t> off[kMaxLevels];
    int64_t npieces = 0;
    int32_t *pieces = nullptr;
    int32_t *pptr = nullptr;
    double *partials = nullptr;  // [npieces]
    unsigned *ticket = nullptr;
    double *S = nullptr;         // per-problem scalars, slot-major: S[slot * count + q]
    int *live = nullptr;         // problem still iterating
    int *niters = nullptr;
    int *nlive = nullptr;
    int grid = 0;                // piece kernels' grid: 0 = one workgroup per piece
    IterGraph graph;             // captured fleet iteration
    double graph_tol = -1.0;
  } fleet;
  // events of CG's phase timers (HPCGX_CG_TIMERS), reused call after call
  struct EventPool {
    std::vector<hipEvent_t> ev;
    ~EventPool() {
      for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
  } tlog;
  // event log of the finest level's multicolour Gauss-Seidel launches
  struct KernelLog {
    bool on = false;
    std::vector<hipEvent_t> ev;  // pairs: start, stop
    size_t used = 0;
    ~KernelLog() {
      for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
  } klog;
};

// Kernel view of a level's 16-bit column ids.  Kernels that use it start
// every wavefront at a multiple of kTile rows, so a wavefront's rows share
// one tile and its bases are wave-uniform (read through the scalar cache).
struct Col16 {
  const uint16_t *off;
  const uint32_t *off8;
  const int32_t *base;   // slot-major bases (Level::cbase)
  const int32_t *tbase;  // tile-major bases (Level::tbase)
  const int32_t *info;
  int64_t ntiles;
  int use8;  // read 8-bit offsets in kTile8 tiles (else their 16-bit ids)
  const uint8_t *amask;  // Level::amask, or null: load every offset word
};
inline Col16 col16_of(const Problem &P, const Level &L) {
  return Col16{L.col16, L.off8, L.cbase, L.tbase, L.tinfo, L.ntiles, P.use_c8 ? 1 : 0,
               P.use_affine ? L.amask : nullptr};
}
// Offset word g of row i on a kTile8 tile: loaded, or lane * 0x01010101
// where the tile's Level::amask bit g says so (amask uniform per tile).
template <bool STREAM>
__device__ inline uint32_t off8_word(const uint32_t *__restrict__ off8, unsigned amask, int g,
                                     int64_t n, int64_t i) {
  if ((amask >> g) & 1u) return (uint32_t)(i % kTile) * 0x01010101u;
  if (STREAM) return __builtin_nontemporal_load(off8 + (int64_t)g * n + i);
  return off8[(int64_t)g * n + i];
}

// Fine row of fine-order transfer row t (Level::ftile).
__device__ inline int64_t fine_row(const int32_t *__restrict__ ftile, int64_t t) {
  return (int64_t)ftile[t / kTile] * kTile + t % kTile;
}

// 16-bit kernels apply to rows [begin, ...) of L when this holds.
inline bool use_col16(const Problem &P, const Level &L, int64_t begin) {
  return P.use_c16 && L.c16 && begin % kTile == 0;
}

// Slots of Problem::scalars used by CG.
enum CgScalar { kRtz = 0, kOldRtz, kPAp, kAlpha, kBeta, kNormr2, kDot, kNumScalars };

// ---- launchers (kernels.hip) ----
void launch_fill(double *x, int64_t n, double v, hipStream_t s);
void launch_copy(const double *x, double *y, int64_t n, hipStream_t s);
void launch_gather(const int32_t *idx, const double *src, double *dst,
                   int64_t n, hipStream_t s);
void launch_spmv(const Problem &P, const Level &L, const double *x, double *y,
                 hipStream_t s);
// With dpart, wavefront k of the launch (rows begin + 64 k ..) also stores
// the sum of r_i * (new x_i) over its rows to dpart[k] (16-bit column id
// path only).
void launch_symgs_rows(const Problem &P, const Level &L, int64_t begin, int64_t end,
                       const double *r, double *x, hipStream_t s, double *dpart = nullptr);
void launch_symgs_list(const Level &L, const int32_t *rows, int64_t count,
                       const double *r, double *x, hipStream_t s);
// The colour step of a forward sweep whose x was zero when the sweep began
// (ComputeMG_ref.cpp:42-47): reads only x of device rows below `begin`
// (rows of colours already updated in this sweep) and takes 0 for every
// other x, so x needs no zero fill; same bits as launch_symgs_rows on a
// zero-filled x.  False (nothing launched) where it does not apply (no
// narrow column ids at begin, no Level::zmask).
bool launch_symgs_rows_zero(const Problem &P, const Level &L, int64_t begin, int64_t end,
                            const double *r, double *x, hipStream_t s);
void launch_restriction(const Level &L, const double *rf, const double *Axf,
                        double *rc, hipStream_t s);
void launch_residual_restriction(const Problem &P, const Level &L, const double *rf,
                                 const double *xf, double *rc, hipStream_t s);
void launch_prolongation(const Level &L, const double *xc, double *xf,
                         hipStream_t s);
void launch_waxpby(int64_t n, double alpha, const double *x, double beta,
                   const double *y, double *w, hipStream_t s);
// result = sum x*y into scalars[slot]; op selects the finishing step
enum DotFinish { kFinishStore = 0, kFinishBeta = 1, kFinishAlpha = 2 };
void launch_dot(Problem &P, int64_t n, const double *x, const double *y,
                double *out, int finish, hipStream_t s);
// The dot's first pass: one workgroup per kDotPer elements, kDotE per
// thread, one partial per workgroup; vectors of up to kDotParts * kDotPer
// elements take it, longer ones the grid-stride reduction.
constexpr int kDotE = 8;
constexpr int64_t kDotPer = (int64_t)kBlock * kDotE;
constexpr int64_t kDotParts = 131072;
// launch_dot with explicit scratch: parts (kDotParts doubles), partials
// (kReduceBlocks doubles), the completion ticket (0 on entry, 0 again on
// exit) and the scalars
void launch_dot_raw(int64_t n, const double *x, const double *y, double *parts,
                    double *partials, unsigned *ticket, double *scalars, double *out, int finish,
                    hipStream_t s);
void launch_zero_u32(unsigned *a, int64_t n, hipStream_t s);
// Sum of part[0..count) in fixed order, finished like launch_dot.
void launch_partials_dot(Problem &P, const double *part, int64_t count, int finish,
                         hipStream_t s);
void launch_spmv_dot(Problem &P, const Level &L, const double *p, double *Ap,
                     hipStream_t s);
// the mixed path's outer Ap = A p, p.Ap from the FP16 off-diagonal values
// and the FP64 diagonal (the FP64 result while those values are exact)
void launch_spmv_dot_half(Problem &P, const Level &L, const double *p, double *Ap,
                          hipStream_t s);
// x == nullptr: r and r.r only (x's step deferred to launch_update_xp)
void launch_update_xr(Problem &P, int64_t n, double *x, double *r,
                      const double *p, const double *Ap, hipStream_t s,
                      float *r32 = nullptr);
// x += alpha p (the previous iteration's step) and p = z + beta p in one pass
void launch_update_xp(int64_t n, double *x, const double *z, double *p,
                      const double *scalars, hipStream_t s);
// x += alpha p alone (the last deferred step)
void launch_step_x(int64_t n, double *x, const double *p, const double *scalars,
                   hipStream_t s);
// mixed precision: dot with, and p update from, the FP32 V-cycle output
void launch_dot32(Problem &P, int64_t n, const double *x, const float *y, double *out,
                  int finish, hipStream_t s);
// stream: non-temporal vector accesses on the fine level (Problem::stream_p)
void launch_update_xp32(int64_t n, double *x, const float *z, double *p,
                        const double *scalars, hipStream_t s);
void launch_update_p32(int64_t n, const float *z, double *p, const double *scalars,
                       hipStream_t s, bool stream);
void launch_update_p(int64_t n, const double *z, double *p,
                     const double *scalars, hipStream_t s, bool stream);
// r = b - A x on the finest level L with r32 = (float)r and r.r (into the
// scalars) in one pass over the row chunks; false (nothing launched) where
// the chunk partials do not cover L
bool launch_residual_fused(Problem &P, const Level &L, const double *b, const double *x,
                           double *r, float *r32, hipStream_t s);
void launch_residual0(Problem &P, int64_t n, const double *b, const double *Ap,
                      double *r, hipStream_t s);
void launch_replace_diag(Level &L, const double *d, hipStream_t s);
void launch_symgs_wave(Problem &P, const Level &L, int dir, const double *r,
                       double *x, hipStream_t s);
int wave_grid_size();

// ---- setup (setup.hip) ----
struct GridPart {  // an nx*ny*nz grid in natural order at rows [off, off + nx*ny*nz)
  int64_t off;
  int nx, ny, nz;
};
struct NaturalLevel {  // device buffers of one level in natural row order
  int nx = 0, ny = 0, nz = 0;
  // The grids the rows come from, when known: they propose the schedule
  // (finalize_level verifies the proposal row by row before using it).
  std::vector<GridPart> parts;
  int64_t n = 0, nnz = 0;
  int32_t *col = nullptr;     // [kSlots * n] column-major
  double *val = nullptr;      // [kSlots * n]
  uint8_t *nnz_row = nullptr; // [n]
  int32_t *f2c = nullptr;     // natural f2c to next level
  int64_t nc = 0;
};
void generate_natural(NaturalLevel &N, hipStream_t s);
// one nx*ny*nz grid at natural rows [off, ...) of N (coarse rows at offc)
void generate_natural_part(NaturalLevel &N, int nx, int ny, int nz, int64_t off,
                           int64_t offc, hipStream_t s);
void finalize_level(Problem &P, int l, NaturalLevel &N, int flags,
                    hipStream_t s);
void link_levels(Problem &P, int l, const NaturalLevel &Nf, hipStream_t s);
void build_wave(Problem &P, Level &L, int slab_factor, int rows, hipStream_t s);
void build_chunks(Problem &P, Level &L, bool interleave);
void drop_wave(Problem &P, Level &L);
void problem_vectors(const Problem &P, double *b, double *x, double *xexact,
                     hipStream_t s);
// The grids level l's rows come from (empty when unknown).
std::vector<GridPart> grid_parts(const Problem &P, int l);
// CheckProblem.cpp:79-148 on the device: every level's matrix and f2c, and
// any of b, x, xexact given (fine level, device order), against the
// generated 27-point problem of the level's grids; throws HPCGX_ECHECK on a
// difference, HPCGX_EINVAL when a level's grid is unknown.  Synchronous.
void check_problem(const Problem &P, const double *b, const double *x, const double *xexact,
                   hipStream_t s);
// ---- host levels (hostload.hip) ----
// N's col / val / nnz_row (allocated for H.nrow rows) from the reference's
// row-pointer storage; throws on a row with 0 or more than 27 entries
void upload_host_level(const hpcgx_host_level &H, NaturalLevel &N, hipStream_t s);

// ---- mixed precision (mixed.hip) ----
struct TempBufD {  // a temporary device buffer freed on scope exit
  void *p = nullptr;
  explicit TempBufD(size_t bytes) { HPCGX_HIP(hipMalloc(&p, bytes ? bytes : 1)); }
  ~TempBufD() { (void)hipFree(p); }
};
void prepare_mixed(Problem &P);
void replace_diag32(Level &L, const double *d, hipStream_t s);
// zero_x: x was zero when the sweep began (the V-cycle's first sweep).
// rz: the backward sweep also leaves the r.z partials of r64 and the new x
// (one per wavefront of rows, colours in backward order).
void symgs32(Problem &P, int l, int ordering, const float *r, float *x,
             hipStream_t s, bool zero_x = false, double *rz = nullptr,
             const double *r64 = nullptr);
void spmv32(const Problem &P, const Level &L, const float *x, float *y, hipStream_t s);
// rz, r64: level l's post-smoother leaves the r.z partials (symgs32)
void mg32(Problem &P, int l, int ordering, const float *r, float *x,
          hipStream_t s, double *rz = nullptr, const double *r64 = nullptr);
// r.z partials of a colour step over `rows` rows: one per wavefront with rows
inline int64_t rz_slots(int64_t rows) { return rows > 0 ? (rows + kTile - 1) / kTile : 0; }
void to_f32(int64_t n, const double *in, float *out, hipStream_t s);
// ---- block PCG (block.hip) ----
void prepare_block(Problem &P, int s);
void mg32_block(Problem &P, int l, int ordering, int s, const float *R, float *X,
                hipStream_t st);
void spmv_block(Problem &P, int l, int s, const double *X, double *Y, hipStream_t st);
void cg_block_mixed(Problem &P, int ordering, int s, const double *B, double *X, int max_iter,
                    double tol, int recompute, int *niters, double *normr, double *normr0,
                    double *times, hipStream_t st);
void to_f64(int64_t n, const float *in, double *out, hipStream_t s);
// ---- fleets of independent problems (fleet.hip) ----
constexpr int kMaxFleet = 1024;  // problems per fleet
constexpr int kPieceRows = 4096; // rows per piece of the fleet's reductions
void prepare_fleet(Problem &P, const std::vector<int> &dims, hipStream_t s);
void cg_fleet(Problem &P, int ordering, const double *b, double *x, int max_iter,
              double tol, bool precond, int flags, int *niters, double *normr,
              double *normr0, double *times, hipStream_t s);
void cg_mixed(Problem &P, int ordering, const double *b, double *x,
              int max_iter, double tol, int recompute, int *niters,
              double *normr, double *normr0, double *times, hipStream_t s);

// ---- solver (solver.hip) ----
// With rz (the buffer of rz_partials), the last sweep of level l's
// smoother also leaves the r.z partials of its rows there.
void symgs(Problem &P, int l, int ordering, const double *r, double *x,
           hipStream_t s, double *rz = nullptr, bool zero_x = false);
bool zero_guess_sweep(const Problem &P, int l, int ordering);
void mg(Problem &P, int l, int ordering, const double *r, double *x,
        hipStream_t s, double *rz = nullptr);
// Number of r.z partials the finest level's colour steps leave, or 0 when
// the multicolour V-cycle in this ordering cannot produce them.
int64_t rz_partials(const Problem &P, int ordering);
void check_status(Problem &P);
void cg(Problem &P, int ordering, const double *b, double *x, int max_iter,
        double tol, bool precond, int flags, int *niters, double *normr,
        double *normr0, double *times, hipStream_t s);

}  // namespace hpcgx
