/*
 * oracle.c -- CPU restatement of the HPCG 3.1 reference (*_ref) PCG path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for hpcg_amd/, never the product).
 * Compiled with -ffp-contract=off so every a*b+c is two roundings, as in the
 * reference built by g++ for x86-64 without FMA.
 *
 * Citations are to the reference source tree (src/...).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* GenerateProblem_ref.cpp:131-183 (single rank: global == local indices,    */
/* SetupHalo_ref.cpp:58-65 copies mtxIndG into mtxIndL).                      */
/* ------------------------------------------------------------------------ */
int64_t or_generate(int nx, int ny, int nz, uint8_t *nnz_row, int32_t *cols,
                    double *vals, uint8_t *diag_slot, double *b, double *x,
                    double *xexact) {
  int64_t total = 0;
  const int64_t nxy = (int64_t)nx * ny;
  for (int iz = 0; iz < nz; iz++) {
    for (int iy = 0; iy < ny; iy++) {
      for (int ix = 0; ix < nx; ix++) {
        const int64_t row = iz * nxy + (int64_t)iy * nx + ix;
        int cnt = 0;
        int32_t *c = cols + OR_SLOTS * row;
        double *v = vals + OR_SLOTS * row;
        /* neighbours in sz, sy, sx order -> ascending column index */
        for (int sz = -1; sz <= 1; sz++) {
          if (iz + sz < 0 || iz + sz >= nz) continue;
          for (int sy = -1; sy <= 1; sy++) {
            if (iy + sy < 0 || iy + sy >= ny) continue;
            for (int sx = -1; sx <= 1; sx++) {
              if (ix + sx < 0 || ix + sx >= nx) continue;
              const int64_t col = row + sz * nxy + (int64_t)sy * nx + sx;
              if (col == row) {
                diag_slot[row] = (uint8_t)cnt;
                v[cnt] = 26.0;
              } else {
                v[cnt] = -1.0;
              }
              c[cnt] = (int32_t)col;
              cnt++;
            }
          }
        }
        /* unused slots are left as (col = row, val = 0) so a full 27-slot
         * traversal adds exact zeros; the kernels below stop at nnz_row. */
        for (int j = cnt; j < OR_SLOTS; j++) {
          c[j] = (int32_t)row;
          v[j] = 0.0;
        }
        nnz_row[row] = (uint8_t)cnt;
        total += cnt;
        if (b) b[row] = 26.0 - ((double)(cnt - 1));
        if (x) x[row] = 0.0;
        if (xexact) xexact[row] = 1.0;
      }
    }
  }
  return total;
}

/* GenerateCoarseProblem.cpp:71-82 */
void or_f2c(int nxf, int nyf, int nzf, int32_t *f2c) {
  const int nxc = nxf / 2, nyc = nyf / 2, nzc = nzf / 2;
  for (int izc = 0; izc < nzc; ++izc)
    for (int iyc = 0; iyc < nyc; ++iyc)
      for (int ixc = 0; ixc < nxc; ++ixc) {
        const int64_t c = (int64_t)izc * nxc * nyc + (int64_t)iyc * nxc + ixc;
        const int64_t f = (int64_t)(2 * izc) * nxf * nyf +
                          (int64_t)(2 * iyc) * nxf + 2 * ixc;
        f2c[c] = (int32_t)f;
      }
}

/* ComputeSPMV_ref.cpp:61-70 */
void or_spmv(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
             const double *vals, const double *x, double *y) {
  for (int64_t i = 0; i < n; i++) {
    double sum = 0.0;
    const int32_t *c = cols + OR_SLOTS * i;
    const double *v = vals + OR_SLOTS * i;
    for (int j = 0; j < nnz_row[i]; j++) sum += v[j] * x[c[j]];
    y[i] = sum;
  }
}

/* One row of ComputeSYMGS_ref.cpp:68-80 (identical in the back sweep, :87-99):
 * subtract every stored term including the diagonal, add the diagonal term
 * back, divide by the diagonal. */
static inline void symgs_row(int64_t i, const uint8_t *nnz_row,
                             const int32_t *cols, const double *vals,
                             const uint8_t *diag_slot, const double *r,
                             double *x) {
  const int32_t *c = cols + OR_SLOTS * i;
  const double *v = vals + OR_SLOTS * i;
  const double d = v[diag_slot[i]];
  double sum = r[i];
  for (int j = 0; j < nnz_row[i]; j++) sum -= v[j] * x[c[j]];
  sum += x[i] * d;
  x[i] = sum / d;
}

/* ComputeSYMGS_ref.cpp:67-100 */
void or_symgs(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
              const double *vals, const uint8_t *diag_slot,
              const int32_t *order, const double *r, double *x) {
  for (int64_t k = 0; k < n; k++)
    symgs_row(order ? order[k] : k, nnz_row, cols, vals, diag_slot, r, x);
  for (int64_t k = n - 1; k >= 0; k--)
    symgs_row(order ? order[k] : k, nnz_row, cols, vals, diag_slot, r, x);
}

/* ComputeRestriction_ref.cpp:51 */
void or_restriction(int64_t nc, const int32_t *f2c, const double *rf,
                    const double *Axf, double *rc) {
  for (int64_t i = 0; i < nc; ++i) rc[i] = rf[f2c[i]] - Axf[f2c[i]];
}

/* ComputeProlongation_ref.cpp:49 */
void or_prolongation(int64_t nc, const int32_t *f2c, const double *xc,
                     double *xf) {
  for (int64_t i = 0; i < nc; ++i) xf[f2c[i]] += xc[i];
}

/* ComputeDotProduct_ref.cpp:51-64 */
double or_dot(int64_t n, const double *x, const double *y) {
  double s = 0.0;
  if (x == y) {
    for (int64_t i = 0; i < n; i++) s += x[i] * x[i];
  } else {
    for (int64_t i = 0; i < n; i++) s += x[i] * y[i];
  }
  return s;
}

/* ComputeWAXPBY_ref.cpp:52-67 */
void or_waxpby(int64_t n, double alpha, const double *x, double beta,
               const double *y, double *w) {
  if (alpha == 1.0) {
    for (int64_t i = 0; i < n; i++) w[i] = x[i] + beta * y[i];
  } else if (beta == 1.0) {
    for (int64_t i = 0; i < n; i++) w[i] = alpha * x[i] + y[i];
  } else {
    for (int64_t i = 0; i < n; i++) w[i] = alpha * x[i] + beta * y[i];
  }
}

/* ComputeResidual.cpp:70-76 */
double or_residual(int64_t n, const double *v1, const double *v2) {
  double r = 0.0;
  for (int64_t i = 0; i < n; i++) {
    const double d = fabs(v1[i] - v2[i]);
    if (d > r) r = d;
  }
  return r;
}

/* ------------------------------------------------------------------------ */
/* Schedules                                                                 */
/* ------------------------------------------------------------------------ */
int or_levels(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
              int32_t *level) {
  int nlev = 0;
  for (int64_t i = 0; i < n; i++) {
    int lv = 0;
    const int32_t *c = cols + OR_SLOTS * i;
    for (int j = 0; j < nnz_row[i]; j++)
      if (c[j] < i && level[c[j]] + 1 > lv) lv = level[c[j]] + 1;
    level[i] = lv;
    if (lv + 1 > nlev) nlev = lv + 1;
  }
  return nlev;
}

/* OptimizeProblem.cpp:43-79 (the colour assignment; the permutation built
 * from it at :81-96 is not restated: its prefix scan at :86-92 is wrong and
 * the result is discarded by the reference). */
int or_colors(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
              int32_t *color) {
  int total = n > 0 ? 1 : 0;
  if (n > 0) color[0] = 0;
  for (int64_t i = 1; i < n; i++) {
    unsigned char used[OR_SLOTS + 1];
    memset(used, 0, sizeof(used));
    const int32_t *c = cols + OR_SLOTS * i;
    for (int j = 0; j < nnz_row[i]; j++)
      if (c[j] < i && color[c[j]] <= OR_SLOTS) used[color[c[j]]] = 1;
    int k = 0;
    while (k < total && used[k]) k++;
    color[i] = k; /* k == total opens a new colour (:72-76) */
    if (k == total) total++;
  }
  return total;
}

/* ------------------------------------------------------------------------ */
/* Hierarchy                                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nx, ny, nz;
  int64_t n, nnz;
  uint8_t *nnz_row, *diag_slot;
  int32_t *cols;
  double *vals;
  int32_t *f2c; /* to the next level (NULL on coarsest) */
  int64_t nc;
  double *Axf, *rc, *xc;
  int32_t *order;
} or_level;

struct or_hier {
  int nlevels;
  or_level lv[8];
  double *b, *x, *xexact;
};

or_hier *or_hier_create(int nx, int ny, int nz, int nlevels) {
  if (nlevels < 1 || nlevels > 8) return NULL;
  or_hier *h = (or_hier *)calloc(1, sizeof(or_hier));
  h->nlevels = nlevels;
  for (int l = 0; l < nlevels; l++) {
    or_level *L = &h->lv[l];
    L->nx = nx; L->ny = ny; L->nz = nz;
    L->n = (int64_t)nx * ny * nz;
    L->nnz_row = (uint8_t *)malloc(L->n);
    L->diag_slot = (uint8_t *)malloc(L->n);
    L->cols = (int32_t *)malloc(sizeof(int32_t) * OR_SLOTS * L->n);
    L->vals = (double *)malloc(sizeof(double) * OR_SLOTS * L->n);
    if (l == 0) {
      h->b = (double *)malloc(sizeof(double) * L->n);
      h->x = (double *)malloc(sizeof(double) * L->n);
      h->xexact = (double *)malloc(sizeof(double) * L->n);
    }
    L->nnz = or_generate(nx, ny, nz, L->nnz_row, L->cols, L->vals,
                         L->diag_slot, l == 0 ? h->b : NULL,
                         l == 0 ? h->x : NULL, l == 0 ? h->xexact : NULL);
    if (l + 1 < nlevels) {
      if (nx % 2 || ny % 2 || nz % 2) { /* GenerateCoarseProblem.cpp:50 */
        or_hier_destroy(h);
        return NULL;
      }
      L->nc = (int64_t)(nx / 2) * (ny / 2) * (nz / 2);
      L->f2c = (int32_t *)malloc(sizeof(int32_t) * L->nc);
      or_f2c(nx, ny, nz, L->f2c);
      L->Axf = (double *)malloc(sizeof(double) * L->n);
      L->rc = (double *)malloc(sizeof(double) * L->nc);
      L->xc = (double *)malloc(sizeof(double) * L->nc);
      nx /= 2; ny /= 2; nz /= 2;
    }
  }
  return h;
}

void or_hier_destroy(or_hier *h) {
  if (!h) return;
  for (int l = 0; l < h->nlevels; l++) {
    or_level *L = &h->lv[l];
    free(L->nnz_row); free(L->diag_slot); free(L->cols); free(L->vals);
    free(L->f2c); free(L->Axf); free(L->rc); free(L->xc); free(L->order);
  }
  free(h->b); free(h->x); free(h->xexact);
  free(h);
}

int or_hier_nlevels(const or_hier *h) { return h->nlevels; }
int64_t or_hier_nrow(const or_hier *h, int l) { return h->lv[l].n; }
int64_t or_hier_nnz(const or_hier *h, int l) { return h->lv[l].nnz; }
void or_hier_dims(const or_hier *h, int l, int *nx, int *ny, int *nz) {
  *nx = h->lv[l].nx; *ny = h->lv[l].ny; *nz = h->lv[l].nz;
}
uint8_t *or_hier_nnz_row(or_hier *h, int l) { return h->lv[l].nnz_row; }
int32_t *or_hier_cols(or_hier *h, int l) { return h->lv[l].cols; }
double *or_hier_vals(or_hier *h, int l) { return h->lv[l].vals; }
uint8_t *or_hier_diag_slot(or_hier *h, int l) { return h->lv[l].diag_slot; }
int32_t *or_hier_f2c(or_hier *h, int l) { return h->lv[l].f2c; }
double *or_hier_b(or_hier *h) { return h->b; }
double *or_hier_x(or_hier *h) { return h->x; }
double *or_hier_xexact(or_hier *h) { return h->xexact; }

int or_hier_set_order(or_hier *h, int l, const int32_t *order) {
  or_level *L = &h->lv[l];
  free(L->order);
  L->order = NULL;
  if (order) {
    L->order = (int32_t *)malloc(sizeof(int32_t) * L->n);
    memcpy(L->order, order, sizeof(int32_t) * L->n);
  }
  return 0;
}

static void level_symgs(or_level *L, const double *r, double *x) {
  or_symgs(L->n, L->nnz_row, L->cols, L->vals, L->diag_slot, L->order, r, x);
}

/* ComputeMG_ref.cpp:39-63 */
int or_mg(or_hier *h, int l, const double *r, double *x) {
  or_level *L = &h->lv[l];
  memset(x, 0, sizeof(double) * L->n); /* :42 ZeroVector(x) */
  if (l + 1 < h->nlevels) {
    level_symgs(L, r, x);                                        /* :47 */
    or_spmv(L->n, L->nnz_row, L->cols, L->vals, x, L->Axf);       /* :49 */
    or_restriction(L->nc, L->f2c, r, L->Axf, L->rc);              /* :51 */
    or_mg(h, l + 1, L->rc, L->xc);                                /* :52 */
    or_prolongation(L->nc, L->f2c, L->xc, x);                     /* :53 */
    level_symgs(L, r, x);                                        /* :55 */
  } else {
    level_symgs(L, r, x);                                        /* :59 */
  }
  return 0;
}

/* CG_ref.cpp:58-144 */
int or_cg(or_hier *h, const double *b, double *x, int max_iter, double tol,
          int precond, int *niters, double *normr, double *normr0,
          double *traj) {
  or_level *A = &h->lv[0];
  const int64_t n = A->n;
  double *r = (double *)malloc(sizeof(double) * n);
  double *z = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n);
  double *Ap = (double *)malloc(sizeof(double) * n);
  double rtz = 0.0, oldrtz = 0.0, alpha = 0.0, beta = 0.0, pAp = 0.0;
  *normr = 0.0;
  *niters = 0;
  memcpy(p, x, sizeof(double) * n);                         /* :87 */
  or_spmv(n, A->nnz_row, A->cols, A->vals, p, Ap);          /* :88 */
  or_waxpby(n, 1.0, b, -1.0, Ap, r);                        /* :89 */
  *normr = sqrt(or_dot(n, r, r));                           /* :90-91 */
  *normr0 = *normr;                                         /* :97 */
  if (traj) traj[0] = 1.0;
  for (int k = 1; k <= max_iter && *normr / *normr0 > tol; k++) {
    if (precond)
      or_mg(h, 0, r, z);                                    /* :104 */
    else
      or_waxpby(n, 1.0, r, 0.0, r, z);                      /* :106 */
    if (k == 1) {
      memcpy(p, z, sizeof(double) * n);                     /* :110 */
      rtz = or_dot(n, r, z);                                /* :111 */
    } else {
      oldrtz = rtz;
      rtz = or_dot(n, r, z);                                /* :114 */
      beta = rtz / oldrtz;
      or_waxpby(n, 1.0, z, beta, p, p);                     /* :116 */
    }
    or_spmv(n, A->nnz_row, A->cols, A->vals, p, Ap);        /* :119 */
    pAp = or_dot(n, p, Ap);                                 /* :120 */
    alpha = rtz / pAp;
    or_waxpby(n, 1.0, x, alpha, p, x);                      /* :122 */
    or_waxpby(n, 1.0, r, -alpha, Ap, r);                    /* :123 */
    *normr = sqrt(or_dot(n, r, r));                         /* :124-125 */
    if (traj) traj[k] = *normr / *normr0;
    *niters = k;
  }
  free(r); free(z); free(p); free(Ap);
  return 0;
}

/* TestCG.cpp:52-123 */
int or_testcg(or_hier *h, double *b, int *out, double *normr_out) {
  or_level *A = &h->lv[0];
  const int64_t n = A->n;
  double *orig_diag = (double *)malloc(sizeof(double) * n);
  double *orig_b = (double *)malloc(sizeof(double) * n);
  double *x = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) {                          /* :62-64 */
    orig_diag[i] = A->vals[OR_SLOTS * i + A->diag_slot[i]];
    orig_b[i] = b[i];
  }
  for (int64_t i = 0; i < n; i++) {                          /* :68-79 */
    const double scale = (i < 9) ? (i + 2) * 1.0e6 : 1.0e6;
    A->vals[OR_SLOTS * i + A->diag_slot[i]] = orig_diag[i] * scale;
    b[i] *= scale;
  }
  int count_pass = 0, count_fail = 0, max_noprec = 0, max_prec = 0;
  int niters = 0;
  double normr = 0.0, normr0 = 0.0;
  for (int k = 0; k < 2; ++k) {                              /* :91-111 */
    const int expected = (k == 1) ? 2 : 12;
    for (int i = 0; i < 2; ++i) {
      memset(x, 0, sizeof(double) * n);
      or_cg(h, b, x, 50, 1.0e-12, k == 1, &niters, &normr, &normr0, NULL);
      if (niters <= expected) ++count_pass; else ++count_fail;
      if (k == 0 && niters > max_noprec) max_noprec = niters;
      if (k == 1 && niters > max_prec) max_prec = niters;
    }
  }
  for (int64_t i = 0; i < n; i++) {                          /* :114-115 */
    A->vals[OR_SLOTS * i + A->diag_slot[i]] = orig_diag[i];
    b[i] = orig_b[i];
  }
  out[0] = count_pass; out[1] = count_fail;
  out[2] = max_noprec; out[3] = max_prec;
  *normr_out = normr;
  free(orig_diag); free(orig_b); free(x);
  return 0;
}

/* TestSymmetry.cpp:60-138 */
int or_testsymmetry(or_hier *h, const double *b, const double *xexact,
                    const double *xr, const double *yr, double *out) {
  or_level *A = &h->lv[0];
  const int64_t n = A->n;
  double *z = (double *)malloc(sizeof(double) * n);
  double *xx = (double *)malloc(sizeof(double) * n);
  const double ANorm = 2 * 26.0;                             /* :80 */
  int count_fail = 0;
  const double yNorm2 = or_dot(n, yr, yr);                   /* :83 */
  or_spmv(n, A->nnz_row, A->cols, A->vals, yr, z);           /* :84 */
  const double xtAy = or_dot(n, xr, z);                      /* :87 */
  const double xNorm2 = or_dot(n, xr, xr);                   /* :91 */
  or_spmv(n, A->nnz_row, A->cols, A->vals, xr, z);           /* :92 */
  const double ytAx = or_dot(n, yr, z);                      /* :95 */
  const double depsym_spmv =
      (double)(fabsl((long double)(xtAy - ytAx)) /
               ((xNorm2 * ANorm * yNorm2 + yNorm2 * ANorm * xNorm2) *
                (DBL_EPSILON)));                             /* :98 */
  if (depsym_spmv > 1.0) ++count_fail;
  or_mg(h, 0, yr, z);                                        /* :105 */
  const double xtMinvy = or_dot(n, xr, z);                   /* :108 */
  or_mg(h, 0, xr, z);                                        /* :112 */
  const double ytMinvx = or_dot(n, yr, z);                   /* :115 */
  const double depsym_mg =
      (double)(fabsl((long double)(xtMinvy - ytMinvx)) /
               ((xNorm2 * ANorm * yNorm2 + yNorm2 * ANorm * xNorm2) *
                (DBL_EPSILON)));                             /* :118 */
  if (depsym_mg > 1.0) ++count_fail;
  memcpy(xx, xexact, sizeof(double) * n);                    /* :122 */
  for (int i = 0; i < 2; ++i) {                              /* :126-132 */
    or_spmv(n, A->nnz_row, A->cols, A->vals, xx, z);
    out[3 + i] = or_residual(n, b, z);
  }
  out[0] = depsym_spmv;
  out[1] = depsym_mg;
  out[2] = count_fail;
  free(z); free(xx);
  return 0;
}

/* TestNorms.cpp:31-46 */
void or_testnorms(int samples, const double *values, double *out) {
  double mean_delta = 0.0;
  for (int i = 0; i < samples; ++i) mean_delta += (values[i] - values[0]);
  const double mean = values[0] + mean_delta / (double)samples;
  double sumdiff = 0.0;
  for (int i = 0; i < samples; ++i)
    sumdiff += (values[i] - mean) * (values[i] - mean);
  const double variance = sumdiff / (double)samples;
  out[0] = mean;
  out[1] = variance;
  out[2] = variance < 1.0e-6 ? 1.0 : 0.0;
}

/* ReportResults.cpp:75-136 (sizeof(double)=8, sizeof(local_int_t)=4).
 * nnz is taken as a 64-bit count (the reference's own counter is int). */
void or_hpcg_model(int nlevels, const int64_t *nrow, const int64_t *nnz,
                   int nsets, int ref_max_iters, int opt_max_iters,
                   double *out) {
  const double fsets = nsets;
  const double fniters = fsets * (double)opt_max_iters;
  const double fnrow = (double)nrow[0];
  const double fnnz = (double)nnz[0];
  const double fnops_ddot = (3.0 * fniters + fsets) * 2.0 * fnrow;      /* :81 */
  const double fnops_waxpby = (3.0 * fniters + fsets) * 2.0 * fnrow;    /* :82 */
  const double fnops_sparsemv = (fniters + fsets) * 2.0 * fnnz;         /* :83 */
  double fnops_precond = 0.0;
  for (int i = 1; i < nlevels; ++i) {                                   /* :87-95 */
    const double f = (double)nnz[i - 1];
    fnops_precond += 1.0 * fniters * 4.0 * f;
    fnops_precond += fniters * 2.0 * f;
    fnops_precond += 1.0 * fniters * 4.0 * f;
  }
  fnops_precond += fniters * 4.0 * ((double)nnz[nlevels - 1]);          /* :97 */
  const double fnops = fnops_ddot + fnops_waxpby + fnops_sparsemv + fnops_precond;
  const double frefnops = fnops * ((double)ref_max_iters) / ((double)opt_max_iters);

  const double D = 8.0, I = 4.0;
  const double fnreads_ddot = (3.0 * fniters + fsets) * 2.0 * fnrow * D;   /* :104 */
  const double fnwrites_ddot = (3.0 * fniters + fsets) * D;                /* :105 */
  const double fnreads_waxpby = (3.0 * fniters + fsets) * 2.0 * fnrow * D; /* :106 */
  const double fnwrites_waxpby = (3.0 * fniters + fsets) * fnrow * D;      /* :107 */
  const double fnreads_sparsemv = (fniters + fsets) * (fnnz * (D + I) + fnrow * D); /* :108 */
  const double fnwrites_sparsemv = (fniters + fsets) * fnrow * D;          /* :110 */
  double fnreads_precond = 0.0, fnwrites_precond = 0.0;
  for (int i = 1; i < nlevels; ++i) {                                      /* :115-127 */
    const double fz = (double)nnz[i - 1], fr = (double)nrow[i - 1];
    fnreads_precond += 1.0 * fniters * (2.0 * fz * (D + I) + fr * D);
    fnwrites_precond += 1.0 * fniters * fr * D;
    fnreads_precond += fniters * (fz * (D + I) + fr * D);
    fnwrites_precond += fniters * fz * D;
    fnreads_precond += 1.0 * fniters * (2.0 * fz * (D + I) + fr * D);
    fnwrites_precond += 1.0 * fniters * fz * D;
  }
  {
    const double fz = (double)nnz[nlevels - 1], fr = (double)nrow[nlevels - 1];
    fnreads_precond += fniters * (2.0 * fz * (D + I) + fr * D);            /* :131 */
    fnwrites_precond += fniters * fr * D;                                  /* :132 */
  }
  const double fnreads = fnreads_ddot + fnreads_waxpby + fnreads_sparsemv + fnreads_precond;
  const double fnwrites = fnwrites_ddot + fnwrites_waxpby + fnwrites_sparsemv + fnwrites_precond;
  out[0] = fnops_ddot;
  out[1] = fnops_waxpby;
  out[2] = fnops_sparsemv;
  out[3] = fnops_precond;
  out[4] = fnops;
  out[5] = frefnops;
  out[6] = fnreads;
  out[7] = fnwrites;
  out[8] = fnreads * ((double)ref_max_iters) / ((double)opt_max_iters);
  out[9] = fnwrites * ((double)ref_max_iters) / ((double)opt_max_iters);
}

/* ------------------------------------------------------------------------ */
/* Mixed precision (see oracle.h)                                            */
/* ------------------------------------------------------------------------ */
/* float -> IEEE binary16 (round to nearest even) -> float, by bit
 * manipulation (no compiler half type needed). */
float or_half_round(double v) {
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  const uint32_t sign = u & 0x80000000u;
  uint32_t a = u & 0x7fffffffu;
  float out;
  if (a >= 0x7f800000u) { /* inf / nan */
    return f;
  }
  if (a >= 0x477ff000u) { /* rounds to >= 65520: overflow to inf */
    uint32_t inf = sign | 0x7f800000u;
    memcpy(&out, &inf, 4);
    return out;
  }
  if (a < 0x33000000u) { /* < 2^-25: rounds to zero */
    uint32_t z = sign;
    memcpy(&out, &z, 4);
    return out;
  }
  int e = (int)(a >> 23) - 127;
  uint32_t h; /* half bits without sign */
  if (e < -14) { /* subnormal half: value = m * 2^-24 */
    const uint32_t mant = (a & 0x7fffffu) | 0x800000u;
    const int shift = -e - 14 + 13; /* bits dropped */
    uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    h = q;
  } else {
    uint32_t q = ((uint32_t)(e + 15) << 10) | ((a & 0x7fffffu) >> 13);
    const uint32_t rem = a & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    h = q;
  }
  /* half bits -> float */
  const uint32_t he = (h >> 10) & 0x1fu, hm = h & 0x3ffu;
  uint32_t r;
  if (he == 0) {
    float m = (float)hm * (1.0f / 16777216.0f); /* hm * 2^-24, exact */
    memcpy(&r, &m, 4);
    r |= sign;
  } else if (he == 31) {
    r = sign | 0x7f800000u;
  } else {
    r = sign | ((he - 15 + 127) << 23) | (hm << 13);
  }
  memcpy(&out, &r, 4);
  return out;
}

static inline void symgs32_row(int64_t i, const uint8_t *nnz_row,
                               const int32_t *cols, const double *vals,
                               const uint8_t *diag_slot, const float *r,
                               float *x) {
  const int32_t *c = cols + OR_SLOTS * i;
  const double *v = vals + OR_SLOTS * i;
  float sum = r[i];
  for (int j = 0; j < nnz_row[i]; j++)
    if (j != diag_slot[i]) sum -= or_half_round(v[j]) * x[c[j]];
  x[i] = sum / (float)v[diag_slot[i]];
}

void or_symgs32(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
                const double *vals, const uint8_t *diag_slot,
                const int32_t *order, const float *r, float *x) {
  for (int64_t k = 0; k < n; k++)
    symgs32_row(order ? order[k] : k, nnz_row, cols, vals, diag_slot, r, x);
  for (int64_t k = n - 1; k >= 0; k--)
    symgs32_row(order ? order[k] : k, nnz_row, cols, vals, diag_slot, r, x);
}

void or_spmv32(int64_t n, const uint8_t *nnz_row, const int32_t *cols,
               const double *vals, const uint8_t *diag_slot, const float *x,
               float *y) {
  for (int64_t i = 0; i < n; i++) {
    const int32_t *c = cols + OR_SLOTS * i;
    const double *v = vals + OR_SLOTS * i;
    float s = 0.0f;
    for (int j = 0; j < nnz_row[i]; j++)
      if (j != diag_slot[i]) s += or_half_round(v[j]) * x[c[j]];
    s += (float)v[diag_slot[i]] * x[i];
    y[i] = s;
  }
}

int or_mg32(or_hier *h, int l, const float *r, float *x) {
  or_level *L = &h->lv[l];
  memset(x, 0, sizeof(float) * L->n);
  if (l + 1 < h->nlevels) {
    or_symgs32(L->n, L->nnz_row, L->cols, L->vals, L->diag_slot, L->order, r, x);
    float *ax = (float *)malloc(sizeof(float) * L->n);
    float *rc = (float *)malloc(sizeof(float) * L->nc);
    float *xc = (float *)malloc(sizeof(float) * L->nc);
    or_spmv32(L->n, L->nnz_row, L->cols, L->vals, L->diag_slot, x, ax);
    for (int64_t c = 0; c < L->nc; c++) rc[c] = r[L->f2c[c]] - ax[L->f2c[c]];
    or_mg32(h, l + 1, rc, xc);
    for (int64_t c = 0; c < L->nc; c++) x[L->f2c[c]] += xc[c];
    or_symgs32(L->n, L->nnz_row, L->cols, L->vals, L->diag_slot, L->order, r, x);
    free(ax); free(rc); free(xc);
  } else {
    or_symgs32(L->n, L->nnz_row, L->cols, L->vals, L->diag_slot, L->order, r, x);
  }
  return 0;
}

int or_cg_mixed(or_hier *h, const double *b, double *x, int max_iter,
                double tol, int recompute, int *niters, double *normr,
                double *normr0, double *traj) {
  or_level *A = &h->lv[0];
  const int64_t n = A->n;
  double *r = (double *)malloc(sizeof(double) * n);
  double *z = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n);
  double *Ap = (double *)malloc(sizeof(double) * n);
  float *r32 = (float *)malloc(sizeof(float) * n);
  float *z32 = (float *)malloc(sizeof(float) * n);
  double rtz = 0.0, oldrtz = 0.0, alpha, beta, pAp;
  *niters = 0;
  memcpy(p, x, sizeof(double) * n);
  or_spmv(n, A->nnz_row, A->cols, A->vals, p, Ap);
  or_waxpby(n, 1.0, b, -1.0, Ap, r);
  *normr = sqrt(or_dot(n, r, r));
  *normr0 = *normr;
  if (traj) traj[0] = 1.0;
  for (int k = 1; k <= max_iter && *normr / *normr0 > tol; k++) {
    for (int64_t i = 0; i < n; i++) r32[i] = (float)r[i];
    or_mg32(h, 0, r32, z32);
    for (int64_t i = 0; i < n; i++) z[i] = (double)z32[i];
    if (k == 1) {
      memcpy(p, z, sizeof(double) * n);
      rtz = or_dot(n, r, z);
    } else {
      oldrtz = rtz;
      rtz = or_dot(n, r, z);
      beta = rtz / oldrtz;
      or_waxpby(n, 1.0, z, beta, p, p);
    }
    or_spmv(n, A->nnz_row, A->cols, A->vals, p, Ap);
    pAp = or_dot(n, p, Ap);
    alpha = rtz / pAp;
    or_waxpby(n, 1.0, x, alpha, p, x);
    if (recompute > 0 && k % recompute == 0) {
      or_spmv(n, A->nnz_row, A->cols, A->vals, x, Ap);
      or_waxpby(n, 1.0, b, -1.0, Ap, r);
    } else {
      or_waxpby(n, 1.0, r, -alpha, Ap, r);
    }
    *normr = sqrt(or_dot(n, r, r));
    if (traj) traj[k] = *normr / *normr0;
    *niters = k;
  }
  free(r); free(z); free(p); free(Ap); free(r32); free(z32);
  return 0;
}
