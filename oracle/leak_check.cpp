// leak_check.cpp -- TEST INFRASTRUCTURE (built by dropin.sh leak, under
// AddressSanitizer): the set-up half of the reference's main.cpp:130-163 --
// geometry, the fine problem, SetupHalo, three coarse levels, the CG work
// vectors -- run through the reference's own generator with the
// hpcg_amd/host headers in place of its SparseMatrix.hpp, Vector.hpp,
// MGData.hpp and CGData.hpp, then torn down the way main.cpp:361-368 does.
// LeakSanitizer fails the run if any row array or vector survives.
#include <fstream>

#include "CGData.hpp"
#include "GenerateCoarseProblem.hpp"
#include "GenerateGeometry.hpp"
#include "GenerateProblem.hpp"
#include "SetupHalo.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"
#include "hpcg.hpp"

int main() {
  HPCG_fout.open("/dev/null");
  const local_int_t n = 16;
  Geometry *geom = new Geometry;
  GenerateGeometry(1, 0, 1, 0, n, n, n, n, n, 0, 0, 0, geom);
  SparseMatrix A;
  InitializeSparseMatrix(A, geom);
  Vector b, x, xexact;
  GenerateProblem(A, &b, &x, &xexact);
  SetupHalo(A);
  SparseMatrix *cur = &A;
  for (int level = 1; level < 4; ++level) {
    GenerateCoarseProblem(*cur);
    cur = cur->Ac;
  }
  CGData data;
  InitializeSparseCGData(A, data);
  Vector x_overlap, b_computed;
  InitializeVector(x_overlap, A.localNumberOfColumns);
  InitializeVector(b_computed, A.localNumberOfRows);

  DeleteMatrix(A);
  DeleteCGData(data);
  DeleteVector(x_overlap);
  DeleteVector(b_computed);
  DeleteVector(x);
  DeleteVector(b);
  DeleteVector(xexact);
  HPCG_fout.close();
  return 0;
}
