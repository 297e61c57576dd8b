/*
 * oracle.h -- CPU restatement of the HPCG 3.1 reference (*_ref) PCG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in hpcg_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  No product code links it.
 *
 * Storage follows the reference's per-row layout (GenerateProblem_ref.cpp:75-124
 * with HPCG_CONTIGUOUS_ARRAYS): row i owns slots [27*i, 27*i+27) of cols/vals,
 * nnz_row[i] of them are used, diag_slot[i] is the slot of the diagonal
 * (the reference keeps a pointer matrixDiagonal[i] into the same storage).
 *
 * Every function cites the reference file:line it follows.  Pinning of each
 * function (golden vectors from the compiled reference, closed forms,
 * brute force) is done by tests/test_oracle.py.
 */
#ifndef HPCG_ORACLE_H
#define HPCG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_SLOTS 27

/* ---- problem generation ------------------------------------------------ */

/* GenerateProblem_ref.cpp:131-183.  Any of b/x/xexact may be NULL.
 * Returns the number of stored nonzeros (64-bit: the reference's int counter
 * wraps at 2^31, GenerateProblem_ref.cpp:126,177). */
int64_t or_generate(int nx, int ny, int nz, uint8_t *nnz_row, int32_t *cols,
                    double *vals, uint8_t *diag_slot, double *b, double *x,
                    double *xexact);

/* GenerateCoarseProblem.cpp:71-82: f2c[coarse row] = fine row of (2ixc,2iyc,2izc). */
void or_f2c(int nxf, int nyf, int nzf, int32_t *f2c);

/* ---- kernels ------------------------------------------------------------ */

/* ComputeSPMV_ref.cpp:61-70 */
void or_spmv(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
             const double *vals, const double *x, double *y);

/* ComputeSYMGS_ref.cpp:67-100.  order==NULL sweeps rows 0..n-1 then n-1..0;
 * otherwise it sweeps order[0..n-1] then order[n-1..0] (a Gauss-Seidel
 * sweep in a permuted row order; same per-row arithmetic). */
void or_symgs(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
              const double *vals, const uint8_t *diag_slot,
              const int32_t *order, const double *r, double *x);

/* ComputeRestriction_ref.cpp:51 */
void or_restriction(int64_t nc, const int32_t *f2c, const double *rf,
                    const double *Axf, double *rc);

/* ComputeProlongation_ref.cpp:49 */
void or_prolongation(int64_t nc, const int32_t *f2c, const double *xc,
                     double *xf);

/* ComputeDotProduct_ref.cpp:51-64 (sequential summation order i=0..n-1) */
double or_dot(int64_t n, const double *x, const double *y);

/* ComputeWAXPBY_ref.cpp:52-67 (including the alpha==1 / beta==1 branches) */
void or_waxpby(int64_t n, double alpha, const double *x, double beta,
               const double *y, double *w);

/* ComputeResidual.cpp:70-76: max_i |v1_i - v2_i| */
double or_residual(int64_t n, const double *v1, const double *v2);

/* ---- schedules (OptimizeProblem hook) ------------------------------------ */

/* Level sets of the lower-triangular dependency DAG of the row ordering:
 * level[i] = 0 if row i has no column j<i, else 1 + max level[j] over those
 * columns.  Rows of one level are independent, so a forward Gauss-Seidel
 * sweep level by level reproduces ComputeSYMGS_ref.cpp:67-82 exactly.
 * Returns the number of levels. */
int or_levels(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
              int32_t *level);

/* Greedy first-fit colouring in natural order, OptimizeProblem.cpp:43-79:
 * colour[i] = smallest colour not used by a column j<i of row i.
 * Returns the number of colours. */
int or_colors(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
              int32_t *color);

/* ---- multigrid hierarchy, MG, CG ---------------------------------------- */

typedef struct or_hier or_hier;

/* main.cpp:133-144: level 0 from GenerateProblem, levels 1..nlevels-1 from
 * GenerateCoarseProblem (2x coarsening per axis), 1 pre / 1 post smoother
 * step (MGData.hpp:51-52 of the reference's InitializeMGData). */
or_hier *or_hier_create(int nx, int ny, int nz, int nlevels);
void or_hier_destroy(or_hier *h);
int or_hier_nlevels(const or_hier *h);
int64_t or_hier_nrow(const or_hier *h, int level);
int64_t or_hier_nnz(const or_hier *h, int level);
void or_hier_dims(const or_hier *h, int level, int *nx, int *ny, int *nz);
uint8_t *or_hier_nnz_row(or_hier *h, int level);
int32_t *or_hier_cols(or_hier *h, int level);
double *or_hier_vals(or_hier *h, int level);
uint8_t *or_hier_diag_slot(or_hier *h, int level);
int32_t *or_hier_f2c(or_hier *h, int level); /* NULL on the coarsest level */
double *or_hier_b(or_hier *h);      /* level 0 only */
double *or_hier_x(or_hier *h);
double *or_hier_xexact(or_hier *h);
/* Gauss-Seidel row order used on a level by or_mg (NULL restores natural
 * order).  The array is copied. */
int or_hier_set_order(or_hier *h, int level, const int32_t *order);

/* ComputeMG_ref.cpp:39-63 applied at `level` of the hierarchy. */
int or_mg(or_hier *h, int level, const double *r, double *x);

/* CG_ref.cpp:58-144.  traj (may be NULL) receives normr/normr0 after
 * iteration k at traj[k], traj[0] = 1.  Returns 0. */
int or_cg(or_hier *h, const double *b, double *x, int max_iter, double tol,
          int precond, int *niters, double *normr, double *normr0,
          double *traj);

/* TestCG.cpp:52-123.  out[0]=count_pass out[1]=count_fail
 * out[2]=niters_max_no_prec out[3]=niters_max_prec; *normr = last normr.
 * b is the level-0 right-hand side (modified and restored). */
int or_testcg(or_hier *h, double *b, int *out, double *normr);

/* TestSymmetry.cpp:60-138 with caller-supplied random vectors x,y
 * (the reference fills them with FillRandomVector, Vector.hpp:84).
 * out[0]=depsym_spmv out[1]=depsym_mg out[2]=count_fail
 * out[3..4]=SpMV residual of the two A*xexact calls against b. */
int or_testsymmetry(or_hier *h, const double *b, const double *xexact,
                    const double *xr, const double *yr, double *out);

/* TestNorms.cpp:31-46: out[0]=mean out[1]=variance out[2]=pass */
void or_testnorms(int samples, const double *values, double *out);

/* ReportResults.cpp:75-136 flop and byte model.  nrow[l], nnz[l] per level.
 * out[0]=ddot out[1]=waxpby out[2]=spmv out[3]=mg out[4]=total flops
 * out[5]=total with convergence overhead, out[6]=reads out[7]=writes
 * (bytes, raw), out[8]=ref-scaled reads, out[9]=ref-scaled writes. */
void or_hpcg_model(int nlevels, const int64_t *nrow, const int64_t *nnz,
                   int nsets, int ref_max_iters, int opt_max_iters,
                   double *out);

/* ---- mixed precision (no counterpart in the reference) -------------------
 * The V-cycle in FP32 with matrix values rounded to FP16 (double -> float ->
 * half, round to nearest even each time) and the diagonal kept separately
 * in FP32.  Row arithmetic of the FP32 kernels:
 *   SYMGS row:   sum = r_i; for each stored j != diag: sum -= h(a_ij) * x_j;
 *                x_i = sum / d_i                  (d_i = (float)a_ii)
 *   A*x row:     s = 0; for each stored j != diag: s += h(a_ij) * x_j;
 *                s += d_i * x_i
 *   restriction: rc = r[f] - (A*x)_f (fused, only at injected rows)
 * These orders are this project's own; they are pinned against the FP64
 * functions and dense algebra (tests/test_oracle.py), not against the
 * reference. */
float or_half_round(double v);
void or_symgs32(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
                const double *vals, const uint8_t *diag_slot,
                const int32_t *order, const float *r, float *x);
void or_spmv32(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
               const double *vals, const uint8_t *diag_slot, const float *x,
               float *y);
/* V-cycle of ComputeMG_ref.cpp:39-63 in FP32 (orders as set on the hierarchy). */
int or_mg32(or_hier *h, int level, const float *r, float *x);
/* CG_ref.cpp:58-144 with M^-1 = the FP32 V-cycle (r rounded to float on the
 * way in, z widened on the way out) and the residual recomputed as
 * r = b - A x in FP64 after every `recompute`-th iteration (0 = never). */
int or_cg_mixed(or_hier *h, const double *b, double *x, int max_iter,
                double tol, int recompute, int *niters, double *normr,
                double *normr0, double *traj);

#ifdef __cplusplus
}
#endif
#endif
