// OptimizedKernels.hpp -- the reference's optimisable C++ surface, served by
// libhpcg_amd on the GPU.  Same signatures, argument meaning and return
// convention (0 = success) as the reference headers cited per function.
#ifndef HPCG_AMD_OPTIMIZEDKERNELS_HPP
#define HPCG_AMD_OPTIMIZEDKERNELS_HPP

#include "CGData.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"

// ComputeSPMV.hpp:20 -- y = A x.
int ComputeSPMV(const SparseMatrix &A, Vector &x, Vector &y);
// ComputeSYMGS.hpp:20 -- one symmetric Gauss-Seidel sweep on x.
int ComputeSYMGS(const SparseMatrix &A, const Vector &r, Vector &x);
// ComputeMG.hpp:20 -- x = M^-1 r, the V-cycle from A's level down.
int ComputeMG(const SparseMatrix &A, const Vector &r, Vector &x);
// ComputeDotProduct.hpp:18-19 -- result = x . y over the first n values.
int ComputeDotProduct(const local_int_t n, const Vector &x, const Vector &y,
                      double &result, double &time_allreduce,
                      bool &isOptimized);
// ComputeWAXPBY.hpp:18-19 -- w = alpha x + beta y.
int ComputeWAXPBY(const local_int_t n, const double alpha, const Vector &x,
                  const double beta, const Vector &y, Vector &w,
                  bool &isOptimized);
// CG.hpp:22-24 -- preconditioned conjugate gradients.
int CG(const SparseMatrix &A, CGData &data, const Vector &b, Vector &x,
       const int max_iter, const double tolerance, int &niters, double &normr,
       double &normr0, double *times, bool doPreconditioning);
// OptimizeProblem.hpp:22 -- builds the device problem.
int OptimizeProblem(SparseMatrix &A, CGData &data, Vector &b, Vector &x,
                    Vector &xexact);
// OptimizeProblem.hpp:28 -- bytes held by the device problem.
double OptimizeProblemMemoryUse(const SparseMatrix &A);

#endif
