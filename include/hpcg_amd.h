/*
 * hpcg_amd.h -- C ABI of the MI355X-native HPCG PCG path (libhpcg_amd.so).
 *
 * Plain pointers and sizes only.  Every vector argument is a DEVICE pointer
 * (hipMalloc'd or a torch tensor's storage) in the problem's device row
 * order unless the name says "natural".  Calls are asynchronous on `stream`
 * (a hipStream_t; NULL = the default stream) unless documented otherwise.
 * Every entry point returns 0 on success and a nonzero HPCGX_E* code on
 * failure (never throws); hpcgx_last_error() describes the last failure of
 * the calling thread.
 *
 * Each declaration names the reference interface (HPCG 3.1, src/...) it
 * replaces.  Calls on one problem handle are single threaded, like the
 * reference (every hot-path function is called from one host thread).
 */
#ifndef HPCG_AMD_H
#define HPCG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPCGX_ABI_VERSION 1

enum {
  HPCGX_OK = 0,
  HPCGX_EINVAL = 1,  /* bad argument (sizes, level, ordering, NULL) */
  HPCGX_EHIP = 2,    /* a HIP runtime call or kernel launch failed */
  HPCGX_ENOMEM = 3,  /* device allocation failed */
  HPCGX_EUNSUP = 4,  /* matrix outside what the setup supports */
  HPCGX_ECHECK = 5   /* hpcgx_check_problem: the problem differs from the
                        generated 27-point problem */
};

/* Gauss-Seidel row orderings for SYMGS / MG / CG */
enum {
  HPCGX_ORDER_LEXICOGRAPHIC = 0, /* the reference's natural row order,
                                    run level by level (level-scheduled) */
  HPCGX_ORDER_MULTICOLOR = 1     /* greedy first-fit colours, one parallel
                                    step per colour */
};

/* flags for problem creation */
enum {
  HPCGX_KEEP_SCHEDULE = 1 /* keep per-row level and colour ids so that
                             hpcgx_get_schedule can return them */
};
/* Order of the rows inside a colour, for a level that is an nx*ny*nz grid in
 * natural order: by bands of `lines` grid lines (y), band by band, so that
 * a colour step can reuse gathered x values while they are cached.  0 = the
 * default, natural order inside each colour (bands measured neutral on
 * MI355X); 1 = natural order as well.  The "wave" tuning option needs
 * natural order.  Results do not depend on it: rows of one colour are
 * independent. */
#define HPCGX_BAND_LINES(lines) (((lines) & 0xff) << 8)
/* Visiting order of the greedy colours in the multicolour sweeps (forward;
 * the backward sweep visits the reverse).  0 = the default: 3, 5, 6, 7, 1,
 * 2, 4, 0, so that colour 0, which on the 27-point grid holds the points
 * the restriction injects, is updated first in the backward sweep and its
 * residual still sees the other colours' updates; 1 = colour id order
 * 0, 1, ..., 7, which leaves that residual at zero and needs more PCG
 * iterations (70 instead of 57 to the 50-iteration lexicographic residual
 * at 512x512x640).  Other values are rejected (HPCGX_EINVAL).  Colours from
 * 8 on follow in id order.  The device row order follows the visiting
 * order. */
#define HPCGX_SWEEP_ORDER(k) (((k) & 0xf) << 20)

typedef struct hpcgx_problem hpcgx_problem; /* device-resident MG hierarchy */

/* Description of one level as the reference stores it (SparseMatrix.hpp:39-77
 * of the reference: arrays of row pointers, <=27 entries per row, columns
 * ascending so lower-triangular terms precede the diagonal). */
typedef struct {
  int nx, ny, nz;                     /* local grid (Geometry.hpp:54-56) */
  int64_t nrow;                       /* localNumberOfRows */
  const char *nonzerosInRow;          /* [nrow] */
  const int *const *mtxIndL;          /* [nrow] -> row's local column ids */
  const double *const *matrixValues;  /* [nrow] -> row's values */
  const double *const *matrixDiagonal;/* [nrow] -> diagonal entry */
  const int *f2cOperator;             /* MGData.hpp:31, NULL on coarsest */
} hpcgx_host_level;

typedef struct {
  int nx, ny, nz;
  int64_t nrow;          /* rows of the level */
  int64_t nnz;           /* stored nonzeros (64-bit) */
  int ncolors;           /* colours of the multicolour schedule */
  int nlevels_sched;     /* hyperplanes of the lexicographic schedule */
  int64_t ncoarse;       /* rows of the next level (0 on coarsest) */
  int wave;              /* 1: multicolour sweeps run as a wavefront */
  int nslabs;            /* slabs of the wavefront schedule */
  int col16;             /* 1: narrow column ids (tile base + offset) built */
  int64_t wide_tiles;    /* tiles of 64 rows that keep 32-bit column ids */
  int64_t byte_tiles;    /* tiles of 64 rows with 8-bit column offsets */
} hpcgx_level_info;

const char *hpcgx_last_error(void);
int hpcgx_abi_version(void);

/* ---- setup --------------------------------------------------------------- */

/* replaces OptimizeProblem (OptimizeProblem.cpp:37-100 of the reference,
 * declared OptimizeProblem.hpp:22): converts every level's CSR to a
 * column-major 27-slot ELL in HBM, builds the lexicographic level schedule
 * and the greedy colour schedule on the device, and reorders rows by colour
 * (in the colours' sweep order, HPCGX_SWEEP_ORDER).
 * `levels[0]` is the fine level; levels[l].f2cOperator maps level l+1 rows to
 * level l rows (GenerateCoarseProblem.cpp:71-82). */
int hpcgx_optimize_problem(const hpcgx_host_level *levels, int nlevels,
                           int flags, hpcgx_problem **out);

/* replaces GenerateProblem + SetupHalo + GenerateCoarseProblem
 * (GenerateProblem_ref.cpp:50-219, SetupHalo_ref.cpp:58-65,
 * GenerateCoarseProblem.cpp:41-111; main.cpp:133-144 builds 4 levels)
 * followed by OptimizeProblem: the 27-point problem on an nx*ny*nz grid is
 * generated directly in device memory, level by level. */
int hpcgx_generate_problem(int nx, int ny, int nz, int nlevels, int flags,
                           hpcgx_problem **out);

/* replaces CheckProblem (CheckProblem.cpp:50-151 of the reference; main.cpp:
 * 150-159 runs it on every level after generation): every level's matrix in
 * HBM -- each row's stored entries, in order, are its 27-point stencil
 * neighbours with value 26 on the diagonal and -1 elsewhere, nonzerosInRow
 * and the diagonal array agree, the nonzero count is prod(3 n_a - 2) per
 * grid -- and each level's f2c (GenerateCoarseProblem.cpp:71-82); and, where
 * given (device pointers, fine level, device order; NULL to skip), b, x and
 * xexact of GenerateProblem_ref.cpp:178-180.  A level's grid is its nx, ny,
 * nz (hpcgx_generate_problem, hpcgx_host_level's nx/ny/nz) or the fleet's
 * grids.  Returns HPCGX_ECHECK naming the level, the first differing natural
 * row and what differs; HPCGX_EINVAL when a level's grid is unknown.
 * hpcgx_generate_problem and hpcgx_generate_fleet run the matrix checks
 * before returning.  Synchronous. */
int hpcgx_check_problem(const hpcgx_problem *p, const double *b, const double *x,
                        const double *xexact, void *stream);

/* replaces DeleteMatrix / DeleteMGData for device state
 * (SparseMatrix.hpp:151-183, MGData.hpp:65-75). */
int hpcgx_problem_destroy(hpcgx_problem *p);

int hpcgx_problem_nlevels(const hpcgx_problem *p);
int hpcgx_problem_info(const hpcgx_problem *p, int level,
                       hpcgx_level_info *info);

/* replaces OptimizeProblemMemoryUse (OptimizeProblem.cpp:103-107): bytes of
 * device memory held for the optimized data structures. */
double hpcgx_optimize_problem_memory_use(const hpcgx_problem *p);

/* b, x, xexact of GenerateProblem_ref.cpp:178-180 (b_i = 27 - nnz_i, x = 0,
 * xexact = 1) written in device order on the fine level. Any may be NULL. */
int hpcgx_problem_vectors(hpcgx_problem *p, double *b, double *x,
                          double *xexact, void *stream);

/* Per-row schedule ids in NATURAL row order, copied to host memory:
 * which = 0 -> level id of the lexicographic schedule, 1 -> colour.
 * Synchronous.  Needs HPCGX_KEEP_SCHEDULE. */
int hpcgx_get_schedule(const hpcgx_problem *p, int level, int which,
                       int32_t *out_host);

/* ---- residency (host touch points of Vector.hpp / SparseMatrix.hpp) ------ */

/* natural -> device order and back (device pointers, length nrow of level). */
int hpcgx_to_device_order(const hpcgx_problem *p, int level,
                          const double *natural, double *devorder,
                          void *stream);
int hpcgx_to_natural_order(const hpcgx_problem *p, int level,
                           const double *devorder, double *natural,
                           void *stream);

/* replaces ReplaceMatrixDiagonal / CopyMatrixDiagonal (SparseMatrix.hpp:
 * 126-145): diagonal in device order (matrix values and the separate
 * diagonal array are both updated). */
int hpcgx_replace_diagonal(hpcgx_problem *p, int level, const double *diag,
                           void *stream);
int hpcgx_copy_diagonal(const hpcgx_problem *p, int level, double *diag,
                        void *stream);

/* ---- kernels --------------------------------------------------------------- */

/* replaces ComputeSPMV (ComputeSPMV.cpp:40-45 -> ComputeSPMV_ref.cpp:47-72):
 * y = A_level * x.  Per-row summation order j = 0..nnz_i-1 as the reference,
 * no fused multiply-add. */
int hpcgx_spmv(const hpcgx_problem *p, int level, const double *x, double *y,
               void *stream);

/* replaces ComputeSYMGS (ComputeSYMGS.cpp:50-55 -> ComputeSYMGS_ref.cpp:
 * 54-103): one forward and one backward Gauss-Seidel sweep on x with right
 * hand side r, in the given ordering (HPCGX_ORDER_*). */
int hpcgx_symgs(const hpcgx_problem *p, int level, int ordering,
                const double *r, double *x, void *stream);

/* replaces ComputeMG (ComputeMG.cpp:33-38 -> ComputeMG_ref.cpp:39-63): the
 * V-cycle from `level` down, x := M^-1 r (x is zeroed first). Uses the
 * problem's own coarse-level vectors. */
int hpcgx_mg(hpcgx_problem *p, int level, int ordering, const double *r,
             double *x, void *stream);

/* replaces ComputeRestriction_ref (ComputeRestriction_ref.cpp:40-54):
 * rc = rf[f2c] - Axf[f2c]; rc has the next level's length. */
int hpcgx_restriction(const hpcgx_problem *p, int level, const double *rf,
                      const double *Axf, double *rc, void *stream);

/* ComputeSPMV_ref + ComputeRestriction_ref fused (ComputeMG_ref.cpp:49-51):
 * rc = rf[f2c] - (A xf)[f2c], computing A xf only at the injected rows. */
int hpcgx_residual_restriction(const hpcgx_problem *p, int level,
                               const double *rf, const double *xf, double *rc,
                               void *stream);

/* replaces ComputeProlongation_ref (ComputeProlongation_ref.cpp:38-52):
 * xf[f2c] += xc. */
int hpcgx_prolongation(const hpcgx_problem *p, int level, const double *xc,
                       double *xf, void *stream);

/* replaces ComputeDotProduct (ComputeDotProduct.cpp:41-47 ->
 * ComputeDotProduct_ref.cpp:46-80): *result_dev = sum_i x_i y_i, summed in a
 * fixed order (bit-identical on every call).  result_dev is device memory. */
int hpcgx_dot(hpcgx_problem *p, int64_t n, const double *x, const double *y,
              double *result_dev, void *stream);

/* As hpcgx_dot but synchronous, result in host memory. */
int hpcgx_dot_host(hpcgx_problem *p, int64_t n, const double *x,
                   const double *y, double *result, void *stream);

/* No counterpart in the reference: hpcgx_dot for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc), with no problem handle.  It runs
 * hpcgx_dot's kernels, so *result_dev is hpcgx_dot's value bit for bit; the
 * reduction scratch (131072 first-pass partials, 4096 finishing partials
 * and a ticket, 1.03 MiB) is a static device array that each GPU holds
 * until the process ends.  Every n >= 0 fits it (vectors longer than
 * 131072 * 2048 elements take the grid-stride reduction, whose 4096
 * workgroups fit the finishing partials); n < 0 gives HPCGX_EINVAL.  It calls no hipMalloc or hipFree; the first call on a
 * device looks up the array's address there (make that call before any
 * capture).  All calls on one GPU share its array: calls that overlap on one
 * GPU race. */
int hpcgx_dot_noalloc(int64_t n, const double *x, const double *y,
                      double *result_dev, void *stream);

/* replaces ComputeWAXPBY (ComputeWAXPBY.cpp:42-48 -> ComputeWAXPBY_ref.cpp:
 * 42-70): w = alpha x + beta y with the reference's alpha==1 and beta==1
 * forms; w may alias x or y. */
int hpcgx_waxpby(int64_t n, double alpha, const double *x, double beta,
                 const double *y, double *w, void *stream);

/* replaces CG (CG.cpp:59-143 of the reference; CG_ref.cpp:58-144 is the same
 * algorithm on the *_ref kernels).  b, x in device order on the fine level;
 * x holds the initial guess and receives the solution.  The work vectors r,
 * z, p, Ap (CGData.hpp:27-32) belong to the problem.  times[0..5] are
 * accumulated like the reference (total, dot, waxpby, spmv, allreduce = 0,
 * MG) when times != NULL; times[1..5] are measured with HIP events only if
 * HPCGX_CG_TIMERS is set in `flags`.  Synchronous. */
enum {
  HPCGX_CG_TIMERS = 1,
  /* run iterations 2.. as replays of one captured hipGraph of the
   * iteration (eager launches when timers, the kernel-timing log or
   * wavefront sweeps are on; with a tolerance the residual is read after
   * each replay).  The graph is kept with the problem and rebuilt when x,
   * the ordering, the preconditioning switch or a tuning knob changes. */
  HPCGX_CG_GRAPH = 2
};
int hpcgx_cg(hpcgx_problem *p, int ordering, const double *b, double *x,
             int max_iter, double tolerance, int precondition, int flags,
             int *niters, double *normr, double *normr0, double *times,
             void *stream);

/* ---- mixed precision -------------------------------------------------------- */
/* No counterpart in the reference (its specification requires FP64
 * throughout; this is a reported variant).  The V-cycle runs in FP32 on FP16
 * matrix values (diagonal slot zeroed) and an FP32 diagonal; vectors of
 * these calls are FP32 device arrays in device order. */

/* Builds the FP16 values, FP32 diagonal and FP32 coarse vectors of every
 * level (also done by the first call below).  HPCGX_EUNSUP if an
 * off-diagonal value overflows FP16. */
int hpcgx_prepare_mixed(hpcgx_problem *p);
/* y = A x with the FP16 values and FP32 diagonal (FP32 accumulation). */
int hpcgx_spmv32(hpcgx_problem *p, int level, const float *x, float *y,
                 void *stream);
/* One symmetric Gauss-Seidel sweep in FP32 (ComputeSYMGS_ref.cpp:67-100's
 * order of rows, FP32 row arithmetic). */
int hpcgx_symgs32(hpcgx_problem *p, int level, int ordering, const float *r,
                  float *x, void *stream);
/* The V-cycle of ComputeMG_ref.cpp:39-63 in FP32. */
int hpcgx_mg32(hpcgx_problem *p, int level, int ordering, const float *r,
               float *x, void *stream);
/* CG.cpp:59-143 in FP64 with the FP32 V-cycle as preconditioner (r rounded
 * to FP32 on the way in, z widened on the way out).  A p uses FP64 products
 * and sums over the FP16 off-diagonal values and the FP64 diagonal (equal
 * to the FP64 SpMV while the off-diagonal values are exact in FP16); after
 * every `recompute`-th iteration r = b - A x is recomputed with the FP64
 * matrix (0 = never).  Iterations 2.. replay one captured hipGraph of the
 * iteration unless the "mixed_graph" knob is 0.  times[0] accumulates the
 * wall time.  Synchronous. */
int hpcgx_cg_mixed(hpcgx_problem *p, int ordering, const double *b, double *x,
                   int max_iter, double tolerance, int recompute, int *niters,
                   double *normr, double *normr0, double *times, void *stream);

/* ---- block PCG over s right-hand sides ----------------------------------------
 * No counterpart in the reference (CG.cpp solves one system).  s independent
 * systems with the same matrix; every kernel reads a matrix row once for all
 * s vectors.  Multivectors are [n][s] row-major (the s values of a row
 * adjacent) in the level's device row order; s must divide 64. */

/* Y = A X (FP64), per right-hand side ComputeSPMV_ref.cpp:62-69's order. */
int hpcgx_spmv_block(hpcgx_problem *p, int level, int s, const double *X, double *Y,
                     void *stream);
/* The FP32 V-cycle of hpcgx_mg32 for s right-hand sides at once. */
int hpcgx_mg32_block(hpcgx_problem *p, int level, int ordering, int s, const float *R,
                     float *X, void *stream);
/* hpcgx_cg_mixed for s right-hand sides: each keeps its own alpha, beta and
 * residual; iterations continue until every system has
 * normr[k]/normr0[k] <= tolerance (or max_iter).  normr and normr0 are host
 * arrays of s doubles.  Synchronous. */
int hpcgx_cg_block_mixed(hpcgx_problem *p, int ordering, int s, const double *B, double *X,
                         int max_iter, double tolerance, int recompute, int *niters,
                         double *normr, double *normr0, double *times, void *stream);

/* ---- fleets of independent problems ----------------------------------------
 * No counterpart in the reference (main.cpp runs one problem per rank).  A
 * fleet holds `count` independent HPCG problems with grids of their own
 * sizes as one handle: level l is the block-diagonal union of the problems'
 * level l, natural rows concatenated problem by problem, so every level
 * entry point above (SpMV, SYMGS, MG, vectors, to/from device order) treats
 * the fleet as one matrix and each launch advances all problems.  Each row's
 * arithmetic is that of its own problem. */

/* Generates problem q on an nx[q] x ny[q] x nz[q] grid (each divisible by
 * 2^(nlevels-1)), q = 0..count-1 (count <= 1024), and sets everything up
 * as hpcgx_generate_problem does.  Synchronous. */
int hpcgx_generate_fleet(int count, const int *nx, const int *ny, const int *nz, int nlevels,
                         int flags, hpcgx_problem **out);
/* *count = problems in the fleet (0 for a single problem).  When
 * row_offsets is not NULL it receives count + 1 entries: problem q owns the
 * natural rows [row_offsets[q], row_offsets[q+1]) of `level`. */
int hpcgx_fleet_info(const hpcgx_problem *p, int level, int *count, int64_t *row_offsets);
/* CG.cpp:59-143 for every problem of a fleet at once, each with its own
 * scalars: problem q iterates while normr[q]/normr0[q] > tolerance, up to
 * max_iter, and then keeps its x.  b and x are fleet vectors (device order).
 * niters, normr and normr0 are host arrays of `count` entries.  flags:
 * HPCGX_CG_GRAPH replays iterations 2.. from one captured graph.  times[0]
 * accumulates the wall time.  Synchronous. */
int hpcgx_cg_fleet(hpcgx_problem *p, int ordering, const double *b, double *x, int max_iter,
                   double tolerance, int precondition, int flags, int *niters, double *normr,
                   double *normr0, double *times, void *stream);

/* ---- tuning ----------------------------------------------------------------- */

/* No counterpart in the reference (tuning knobs of this implementation):
 *   "wave"             1/0: multicolour sweeps of large levels as one
 *                      wavefront launch per direction (default 0: on
 *                      MI355X it measured slower than one launch per
 *                      colour step)
 *   "wave_rows"        rows per thread of a wavefront task (1..16)
 *   "wave_min_rows"    levels with at least this many rows get a wavefront
 *                      schedule (default 2^23)
 *   "wave_slab_factor" slab width in units of the matrix bandwidth (>= 1)
 *   "xcd_remap"        1/0: map workgroups to XCD-contiguous row ranges
 *   "pair_rows"        1/0: colour steps with two adjacent rows per thread
 *                      (16-byte matrix loads)
 *   "stream_min_bytes" matrices larger than this are read with non-temporal
 *                      loads (default 512e6)
 *   "col16"            1/0: read narrow column ids (a per-tile base plus an
 *                      8- or 16-bit offset) where a level has them
 *                      (default 1)
 *   "col8"             1/0: use the 8-bit offsets of tiles that have them
 *                      (else their 16-bit ones; default 1)
 *   "mixed_graph"      1/0: hpcgx_cg_mixed and hpcgx_cg_block_mixed replay a
 *                      captured iteration (default 1)
 *   "stream_p"         1/0: the CG p update reads and writes the fine-level
 *                      vectors with non-temporal accesses (default 0)
 *   "fp16_pairs"       1/0: the FP32 V-cycle's colour steps on 8-bit tiles
 *                      read FP16 values two slots per word beside the 8-bit
 *                      offsets (3 bytes per slot; else the 4-byte packed
 *                      value + 16-bit column word; default 1)
 *   "affine8"          1/0: skip the 8-bit offset words that setup found to
 *                      equal the row's position in its tile for every row of
 *                      the tile (each of the word's slots then reads the
 *                      slot base plus that position): the FP32 V-cycle per
 *                      word, the FP64 kernels on tiles whose words all are
 *                      (default 1)
 *   "rz_fused"         1/0: the FP64 and the mixed-precision CG take r.z
 *                      from partials that the V-cycle's last finest-level
 *                      colour sweep leaves (each row's r and final z are at
 *                      hand there), summed by a second launch, instead of a
 *                      dot kernel that reads r and z again (default 1;
 *                      applies to the multicolour colour steps with narrow
 *                      column ids, FP16 value pairs in the mixed V-cycle).
 *                      Changes the dot's summation order.
 *   "zero_guess"       1/0: the V-cycle's first forward sweep of a level
 *                      (multicolour ordering, narrow column ids; FP64, FP32
 *                      and block V-cycles) starts from the zero x without
 *                      filling it: each colour step reads only the slots
 *                      and x values of colours already updated in that
 *                      sweep (default 1; the same x as a zero fill followed
 *                      by the full sweep)
 *   "block_nt"         1/0: the block PCG's vector kernels (r.z, the p and
 *                      r updates) use non-temporal loads and stores on
 *                      multivectors of 2^24 or more values (default 1; the
 *                      same bits)
 *   "resid_fused"      1/0: hpcgx_cg_mixed's (and, for s = 4, 8, 16,
 *                      hpcgx_cg_block_mixed's) FP64 residual recompute
 *                      (r = b - A x, its FP32 copy and r.r) in one pass over
 *                      the row chunks, or as SpMV, residual and rounding
 *                      passes (default 1; the same r, r.r summed in another
 *                      order)
 *   "hc_pairs"         1/0: hpcgx_cg_mixed's FP64 A p reads the 8-bit
 *                      column tiles as FP16 value pairs and 8-bit offsets
 *                      (those of the FP32 V-cycle) rather than 4-byte
 *                      offset/value words, and its one-pass residual
 *                      recompute ("resid_fused") reads the FP16 values and
 *                      the FP64 diagonal rather than the FP64 matrix, where
 *                      FP16 holds every off-diagonal value exactly
 *                      (default 1; the same bits)
 *   "zero_live"        0..15: the zero-x colour steps (see "zero_guess")
 *                      issue loads for the live slots only, eight per chunk,
 *                      instead of every slot through empty views; bit 0 the
 *                      FP64 finest level, bit 1 the FP64 coarser levels,
 *                      bit 2 the FP32 V-cycle, bit 3 the block V-cycle
 *                      (default 11; the same x)
 *   "defer_x"          1/0: hpcgx_cg, hpcgx_cg_mixed and (where "xr_rows"
 *                      applies) hpcgx_cg_block_mixed take each
 *                      iteration's x += alpha p with the next iteration's p
 *                      update, in one pass over x and p, and the last one
 *                      after the loop (the mixed CG also before each FP64
 *                      residual recompute) (default 1; the same x and
 *                      residuals)
 *   "gs_block"         threads per workgroup of the finest level's FP64
 *                      colour steps: 64, 128, 256 or 512 (default 256)
 *   "gs_waves"         1/0: the finest level's FP64 colour steps run under
 *                      register caps that allow 7 waves per SIMD (default
 *                      1); 0 leaves the register count to the compiler
 *   "x_pairs"          0-15: on tiles whose offset words are all affine and
 *                      whose slot 3m + 2 has the base one above slot 3m's
 *                      (all m), the finest level's kernels take slot 3m +
 *                      2's x values from slot 3m's moved one lane
 *                      (cross-lane permute) instead of a second gather.
 *                      Bit 0: the FP64 colour steps without r.z partials
 *                      (at 256 threads per workgroup with the register
 *                      caps), the chunked SpMV + dot and the fused residual
 *                      and restriction; bit 1: the FP64 colour step that
 *                      leaves r.z partials; bit 2: the FP32 V-cycle's colour
 *                      steps on FP16 value pairs; bit 3: the FP64 colour
 *                      steps of the coarser levels (default 3; the same bits)
 *   "block_pairs"      1/0: the block PCG's finest colour steps at s = 8,
 *                      on tiles where "x_pairs" applies, take slot 3m + 2's
 *                      neighbour rows from slot 3m's moved one row (2
 *                      lanes) instead of a second gather (default 0; the
 *                      same bits)
 *   "xr_rows"          1/0: CG's r update (with x's step where it is not
 *                      deferred) runs one row per thread and sums r.r from
 *                      per-workgroup partials in a second launch, instead
 *                      of a grid-stride kernel that finishes r.r in its
 *                      last workgroup (default 1).  Changes r.r's summation
 *                      order.  In the block PCG (s dividing 64) it also
 *                      selects the form for r_k.z_k and the r update, four
 *                      elements per thread, which changes the order of
 *                      those sums.
 *   "spmv_dot_blocks"  workgroups of the persistent SpMV + dot kernel
 *                      (8..8192, rounded down to a multiple of 8; default
 *                      4096).  Changes the dot's summation order.
 *   "spmv_dot_chunks"  1/0: the fused SpMV + dot of CG, of the mixed CG and
 *                      of the block PCG (s = 4, 8, 16) runs one workgroup
 *                      per row chunk (s/2 per chunk in the block PCG), and
 *                      a second launch sums the partials in a fixed order
 *                      (default 1; 0: the persistent kernels, whose
 *                      workgroups sweep the rows in strides).  Changes the
 *                      dots' summation order.
 *   "spmv_order"       0: the SpMV kernels visit row chunks interleaved over
 *                      the colour blocks (default); 1: colour block by
 *                      colour block.  Changes the dots' summation order.
 *   "fleet_blocks"     workgroups of a fleet's persistent CG vector kernels
 *                      (1..65536), or 0: one workgroup per piece, each
 *                      reduction finished by a one-workgroup launch of its
 *                      own (default 0; fleets only; the same bits)
 * Synchronous.  Unknown keys return HPCGX_EINVAL. */
int hpcgx_tune(hpcgx_problem *p, const char *key, double value);

/* No counterpart in the reference: picks the tuning knobs for this problem
 * and machine by timing `iterations` PCG iterations (plus one warm-up) per
 * candidate, one knob at a time ("col8", "spmv_dot_chunks", "spmv_dot_blocks", "spmv_order",
 * "xcd_remap", and, where the finest level has narrow column ids, "gs_block", "gs_waves" and
 * "x_pairs"; a fleet times hpcgx_cg_fleet and tunes "fleet_blocks" in place of
 * "spmv_dot_blocks"), keeping the fastest value of each.  Each knob's candidates are timed
 * in two interleaved rounds and keep their faster time.  Writes one line per candidate (that
 * time) and per choice into `report` (may be NULL).  Uses the default stream; synchronous.
 * Changing "spmv_dot_blocks" changes the dots' summation order. */
int hpcgx_autotune(hpcgx_problem *p, int ordering, int iterations, char *report,
                   int report_len);

/* No counterpart in the reference: waits for `stream` and reports a
 * wavefront sweep whose dependency wait gave up (HPCGX_EHIP).  hpcgx_cg
 * performs the same check before returning. */
int hpcgx_check_status(hpcgx_problem *p, void *stream);

/* ---- measurement ----------------------------------------------------------- */

/* No counterpart in the reference (its timers are host clocks around whole
 * kernels, CG.cpp:36-37): records a HIP event pair on the launch stream
 * around every launch of the finest level's multicolour Gauss-Seidel step,
 * the kernel that dominates the V-cycle.  enable = 1 clears the log and
 * starts recording, 0 stops. */
int hpcgx_kernel_timing(hpcgx_problem *p, int enable);
/* Synchronous: summed event time (ms) and number of recorded launches. */
int hpcgx_kernel_timing_read(hpcgx_problem *p, double *total_ms,
                             int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* HPCG_AMD_H */
