// ReferencePhases.cpp -- the kernels the reference's main.cpp runs before
// OptimizeProblem (the SpMV+MG timing phase, main.cpp:176-190, and the
// 50-iteration reference CG that fixes the tolerance, :214-238).  Here they
// run on the device problem in the reference's natural Gauss-Seidel order
// (the level-scheduled sweep), which reproduces ComputeSYMGS_ref and
// ComputeMG_ref bit for bit; the reference CG differs from CG_ref only in
// the summation order of its dot products.
#include "../DeviceProblem.hpp"
#include "../OptimizedKernels.hpp"
#include "Driver.hpp"

using namespace hpcg_amd;

int ComputeSPMV_ref(const SparseMatrix &A, Vector &x, Vector &y) {
  return ComputeSPMV(A, x, y);  // one SpMV kernel: same bits in any row order
}

int ComputeMG_ref(const SparseMatrix &A, const Vector &r, Vector &x) {
  return DeviceMG(A, r, x, HPCGX_ORDER_LEXICOGRAPHIC);
}

int CG_ref(const SparseMatrix &A, CGData &data, const Vector &b, Vector &x, const int max_iter,
           const double tolerance, int &niters, double &normr, double &normr0, double *times,
           bool doPreconditioning) {
  (void)data;
  return DeviceCG(A, b, x, max_iter, tolerance, niters, normr, normr0, times,
                  doPreconditioning, HPCGX_ORDER_LEXICOGRAPHIC);
}
