// CGData.hpp -- CG work vectors (the reference's CGData.hpp:27-63).  The
// optimized CG keeps its work vectors inside the device problem; these host
// vectors are allocated for interface compatibility.
#ifndef HPCG_AMD_CGDATA_HPP
#define HPCG_AMD_CGDATA_HPP

#include "SparseMatrix.hpp"
#include "Vector.hpp"

struct CGData_STRUCT {
  Vector r;   // residual
  Vector z;   // preconditioned residual
  Vector p;   // direction
  Vector Ap;  // A * direction
};
typedef struct CGData_STRUCT CGData;

inline void InitializeSparseCGData(SparseMatrix &A, CGData &data) {
  InitializeVector(data.r, A.localNumberOfRows);
  InitializeVector(data.z, A.localNumberOfColumns);
  InitializeVector(data.p, A.localNumberOfColumns);
  InitializeVector(data.Ap, A.localNumberOfRows);
}

inline void DeleteCGData(CGData &data) {
  DeleteVector(data.r);
  DeleteVector(data.z);
  DeleteVector(data.p);
  DeleteVector(data.Ap);
}

#endif
