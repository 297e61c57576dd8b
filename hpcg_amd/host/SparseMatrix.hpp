// SparseMatrix.hpp -- the HPCG sparse matrix, with its optimized device copy.
//
// Same struct fields and helper names as the reference's SparseMatrix.hpp:
// 39-183 (single rank: the MPI-only members are left out, as the reference
// does under HPCG_NO_MPI).  After OptimizeProblem, `optimizationData` points
// at the level's place in the device problem (DeviceProblem.hpp);
// ReplaceMatrixDiagonal then updates the device copy too, and DeleteMatrix
// releases it.  The host CSR stays valid and in step with the device.
#ifndef HPCG_AMD_SPARSEMATRIX_HPP
#define HPCG_AMD_SPARSEMATRIX_HPP

#include <cassert>
#include <unordered_map>
#include <vector>

#include "Geometry.hpp"
#include "MGData.hpp"
#include "Vector.hpp"

using GlobalToLocalMap = std::unordered_map<global_int_t, local_int_t>;

struct SparseMatrix_STRUCT {
  char *title;
  Geometry *geom;
  global_int_t totalNumberOfRows;
  global_int_t totalNumberOfNonzeros;
  local_int_t localNumberOfRows;
  local_int_t localNumberOfColumns;
  local_int_t localNumberOfNonzeros;
  char *nonzerosInRow;        // <= 27 per row
  global_int_t **mtxIndG;     // global column ids per row
  local_int_t **mtxIndL;      // local column ids per row
  double **matrixValues;      // values per row
  double **matrixDiagonal;    // pointer to each row's diagonal value
  GlobalToLocalMap globalToLocalMap;
  std::vector<global_int_t> localToGlobalMap;
  mutable bool isDotProductOptimized;
  mutable bool isSpmvOptimized;
  mutable bool isMgOptimized;
  mutable bool isWaxpbyOptimized;
  mutable struct SparseMatrix_STRUCT *Ac;  // next coarser level
  mutable MGData *mgData;                   // transfer data to Ac
  void *optimizationData;                   // device level (MatrixResidency.cpp)
};
typedef struct SparseMatrix_STRUCT SparseMatrix;

// device hooks (DeviceProblem.cpp); no-ops before OptimizeProblem
void MatrixReplaceDiagonalOnDevice(SparseMatrix &A, const Vector &diagonal);
void MatrixReleaseDevice(SparseMatrix &A);

inline void InitializeSparseMatrix(SparseMatrix &A, Geometry *geom) {
  A.title = 0;
  A.geom = geom;
  A.totalNumberOfRows = 0;
  A.totalNumberOfNonzeros = 0;
  A.localNumberOfRows = 0;
  A.localNumberOfColumns = 0;
  A.localNumberOfNonzeros = 0;
  A.nonzerosInRow = 0;
  A.mtxIndG = 0;
  A.mtxIndL = 0;
  A.matrixValues = 0;
  A.matrixDiagonal = 0;
  // flags stay true: every kernel of this build is the optimized one
  A.isDotProductOptimized = true;
  A.isSpmvOptimized = true;
  A.isMgOptimized = true;
  A.isWaxpbyOptimized = true;
  A.mgData = 0;
  A.Ac = 0;
  A.optimizationData = 0;
}

inline void CopyMatrixDiagonal(SparseMatrix &A, Vector &diagonal) {
  assert(A.localNumberOfRows == diagonal.localLength);
  for (local_int_t i = 0; i < A.localNumberOfRows; ++i)
    diagonal.values[i] = *(A.matrixDiagonal[i]);
}

inline void ReplaceMatrixDiagonal(SparseMatrix &A, Vector &diagonal) {
  assert(A.localNumberOfRows == diagonal.localLength);
  for (local_int_t i = 0; i < A.localNumberOfRows; ++i)
    *(A.matrixDiagonal[i]) = diagonal.values[i];
  MatrixReplaceDiagonalOnDevice(A, diagonal);
}

// The row arrays are freed the way the generator allocated them
// (GenerateProblem_ref.cpp:103-117 of the reference): one array per row by
// default, one block per array under HPCG_CONTIGUOUS_ARRAYS (the xhpcg_amd
// driver's generator is built that way).
inline void DeleteMatrix(SparseMatrix &A) {
  MatrixReleaseDevice(A);
#ifndef HPCG_CONTIGUOUS_ARRAYS
  for (local_int_t i = 0; i < A.localNumberOfRows; ++i) {
    if (A.matrixValues) delete[] A.matrixValues[i];
    if (A.mtxIndG) delete[] A.mtxIndG[i];
    if (A.mtxIndL) delete[] A.mtxIndL[i];
  }
#else
  if (A.localNumberOfRows > 0) {
    if (A.matrixValues) delete[] A.matrixValues[0];
    if (A.mtxIndG) delete[] A.mtxIndG[0];
    if (A.mtxIndL) delete[] A.mtxIndL[0];
  }
#endif
  delete[] A.title;
  delete[] A.nonzerosInRow;
  delete[] A.mtxIndG;
  delete[] A.mtxIndL;
  delete[] A.matrixValues;
  delete[] A.matrixDiagonal;
  A.title = 0;
  A.nonzerosInRow = 0;
  A.mtxIndG = 0;
  A.mtxIndL = 0;
  A.matrixValues = 0;
  A.matrixDiagonal = 0;
  if (A.geom) {
    DeleteGeometry(*A.geom);
    delete A.geom;
    A.geom = 0;
  }
  if (A.Ac) {
    DeleteMatrix(*A.Ac);
    delete A.Ac;
    A.Ac = 0;
  }
  if (A.mgData) {
    DeleteMGData(*A.mgData);
    delete A.mgData;
    A.mgData = 0;
  }
}

#endif
