// Validate.cpp -- the reference's validation phase, run against the
// optimized (GPU) functions: TestCG.cpp:52-123, TestSymmetry.cpp:60-138,
// TestNorms.cpp:31-46, ComputeResidual.cpp:49-89.  Same steps, thresholds
// and pass/fail counting.
#include <cfloat>
#include <cmath>
#include <cstdio>

#include "../OptimizedKernels.hpp"
#include "Driver.hpp"

int TestCG(SparseMatrix &A, CGData &data, Vector &b, Vector &x, TestCGData &t) {
  double times[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  Vector origDiagA, exaggeratedDiagA, origB;
  InitializeVector(origDiagA, A.localNumberOfRows);
  InitializeVector(exaggeratedDiagA, A.localNumberOfRows);
  InitializeVector(origB, A.localNumberOfRows);
  CopyMatrixDiagonal(A, origDiagA);
  CopyVector(origDiagA, exaggeratedDiagA);
  CopyVector(b, origB);

  // a strongly dominant diagonal: CG converges in about 10 iterations
  // whatever the size; the first 9 global rows get distinct scales
  for (local_int_t i = 0; i < A.localNumberOfRows; ++i) {
    const global_int_t gid = A.localToGlobalMap[i];
    const double scale = gid < 9 ? (gid + 2) * 1.0e6 : 1.0e6;
    ScaleVectorValue(exaggeratedDiagA, i, scale);
    ScaleVectorValue(b, i, scale);
  }
  ReplaceMatrixDiagonal(A, exaggeratedDiagA);

  const int maxIters = 50, numberOfCgCalls = 2;
  const double tolerance = 1.0e-12;
  t.expected_niters_no_prec = 12;
  t.expected_niters_prec = 2;
  t.niters_max_no_prec = 0;
  t.niters_max_prec = 0;
  t.count_pass = t.count_fail = 0;
  for (int k = 0; k < 2; ++k) {
    const bool prec = k == 1;
    const int expected = prec ? t.expected_niters_prec : t.expected_niters_no_prec;
    for (int i = 0; i < numberOfCgCalls; ++i) {
      ZeroVector(x);
      int niters = 0;
      double normr = 0, normr0 = 0;
      int ierr = CG(A, data, b, x, maxIters, tolerance, niters, normr, normr0, times, prec);
      if (ierr) std::printf("Error in call to CG: %d.\n", ierr);
      if (niters <= expected) {
        ++t.count_pass;
      } else {
        ++t.count_fail;
        std::printf("TestCG: expected %d iterations, performed %d\n", expected, niters);
      }
      int &mx = prec ? t.niters_max_prec : t.niters_max_no_prec;
      if (niters > mx) mx = niters;
      t.normr = normr;
      std::printf("TestCG call [%d] preconditioned [%d] iterations [%d] scaled residual [%g]\n",
                  i, (int)prec, niters, normr / normr0);
    }
  }

  ReplaceMatrixDiagonal(A, origDiagA);
  CopyVector(origB, b);
  DeleteVector(origDiagA);
  DeleteVector(exaggeratedDiagA);
  DeleteVector(origB);
  return 0;
}

int TestSymmetry(SparseMatrix &A, Vector &b, Vector &xexact, TestSymmetryData &t) {
  const local_int_t nrow = A.localNumberOfRows, ncol = A.localNumberOfColumns;
  Vector x_ncol, y_ncol, z_ncol;
  InitializeVector(x_ncol, ncol);
  InitializeVector(y_ncol, ncol);
  InitializeVector(z_ncol, ncol);
  double t4 = 0.0;
  t.count_fail = 0;
  FillRandomVector(x_ncol);
  FillRandomVector(y_ncol);
  const double ANorm = 2 * 26.0;  // bound on |A| used for the scaling
  double xNorm2 = 0, yNorm2 = 0, xtAy = 0, ytAx = 0, xtMinvy = 0, ytMinvx = 0;
  int ierr = 0;
  auto check = [&](const char *what) {
    if (ierr) std::printf("Error in call to %s: %d.\n", what, ierr);
  };

  // x'Ay against y'Ax
  ierr = ComputeDotProduct(nrow, y_ncol, y_ncol, yNorm2, t4, A.isDotProductOptimized); check("dot");
  ierr = ComputeSPMV(A, y_ncol, z_ncol); check("SpMV");
  ierr = ComputeDotProduct(nrow, x_ncol, z_ncol, xtAy, t4, A.isDotProductOptimized); check("dot");
  ierr = ComputeDotProduct(nrow, x_ncol, x_ncol, xNorm2, t4, A.isDotProductOptimized); check("dot");
  ierr = ComputeSPMV(A, x_ncol, z_ncol); check("SpMV");
  ierr = ComputeDotProduct(nrow, y_ncol, z_ncol, ytAx, t4, A.isDotProductOptimized); check("dot");
  const double denom = (xNorm2 * ANorm * yNorm2 + yNorm2 * ANorm * xNorm2) * DBL_EPSILON;
  t.depsym_spmv = (double)(std::fabs((long double)(xtAy - ytAx)) / denom);
  if (t.depsym_spmv > 1.0) ++t.count_fail;
  std::printf("Departure from symmetry (scaled) for SpMV abs(x'*A*y - y'*A*x) = %g\n",
              t.depsym_spmv);

  // x'M^-1 y against y'M^-1 x
  ierr = ComputeMG(A, y_ncol, z_ncol); check("MG");
  ierr = ComputeDotProduct(nrow, x_ncol, z_ncol, xtMinvy, t4, A.isDotProductOptimized); check("dot");
  ierr = ComputeMG(A, x_ncol, z_ncol); check("MG");
  ierr = ComputeDotProduct(nrow, y_ncol, z_ncol, ytMinvx, t4, A.isDotProductOptimized); check("dot");
  t.depsym_mg = (double)(std::fabs((long double)(xtMinvy - ytMinvx)) / denom);
  if (t.depsym_mg > 1.0) ++t.count_fail;
  std::printf("Departure from symmetry (scaled) for MG abs(x'*Minv*y - y'*Minv*x) = %g\n",
              t.depsym_mg);

  // A xexact reproduces b
  CopyVector(xexact, x_ncol);
  for (int i = 0; i < 2; ++i) {
    double residual = 0.0;
    ierr = ComputeSPMV(A, x_ncol, z_ncol); check("SpMV");
    ierr = ComputeResidual(nrow, b, z_ncol, residual); check("compute_residual");
    t.spmv_residual[i] = residual;
    std::printf("SpMV call [%d] Residual [%g]\n", i, residual);
  }
  DeleteVector(x_ncol);
  DeleteVector(y_ncol);
  DeleteVector(z_ncol);
  return 0;
}

int TestNorms(TestNormsData &t) {
  double mean_delta = 0.0;
  for (int i = 0; i < t.samples; ++i) mean_delta += t.values[i] - t.values[0];
  t.mean = t.values[0] + mean_delta / (double)t.samples;
  double sumdiff = 0.0;
  for (int i = 0; i < t.samples; ++i) sumdiff += (t.values[i] - t.mean) * (t.values[i] - t.mean);
  t.variance = sumdiff / (double)t.samples;
  t.pass = t.variance < 1.0e-6;
  return 0;
}

int ComputeResidual(const local_int_t n, const Vector &v1, const Vector &v2, double &residual) {
  double r = 0.0;
#pragma omp parallel for reduction(max : r)
  for (local_int_t i = 0; i < n; ++i) {
    const double d = std::fabs(v1.values[i] - v2.values[i]);
    if (d > r) r = d;
  }
  residual = r;
  return 0;
}
