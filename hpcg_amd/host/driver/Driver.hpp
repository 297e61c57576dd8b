// Driver.hpp -- the benchmark driver's own pieces: problem generation,
// problem checks, the reference-ordering phases, the validation tests and
// the report.  Structs and function names follow the reference's headers
// (cited per item) so that the driver reads like its main.cpp.
#ifndef HPCG_AMD_DRIVER_HPP
#define HPCG_AMD_DRIVER_HPP

#include <string>

#include "../CGData.hpp"
#include "../Geometry.hpp"
#include "../SparseMatrix.hpp"
#include "../Vector.hpp"

// ---- setup (GenerateProblem.hpp:20, GenerateCoarseProblem.hpp:19,
//      CheckProblem.hpp:20, CheckAspectRatio.hpp) ------------------------
void GenerateProblem(SparseMatrix &A, Vector *b, Vector *x, Vector *xexact);
void GenerateCoarseProblem(const SparseMatrix &A);
int CheckProblem(const SparseMatrix &A, const Vector *b, const Vector *x,
                 const Vector *xexact);
int CheckAspectRatio(double smallest_ratio, int x, int y, int z,
                     const char *what);

// ---- phases the reference runs with its *_ref kernels before
//      OptimizeProblem (main.cpp:176-238): the same operations on the device
//      in the reference's natural Gauss-Seidel order ---------------------
int ComputeSPMV_ref(const SparseMatrix &A, Vector &x, Vector &y);
int ComputeMG_ref(const SparseMatrix &A, const Vector &r, Vector &x);
int CG_ref(const SparseMatrix &A, CGData &data, const Vector &b, Vector &x,
           const int max_iter, const double tolerance, int &niters,
           double &normr, double &normr0, double *times,
           bool doPreconditioning);

// ---- validation (TestCG.hpp:30-41, TestSymmetry.hpp:28-35,
//      TestNorms.hpp:25-34, ComputeResidual.hpp:18) ---------------------
struct TestCGData {
  int count_pass;
  int count_fail;
  int expected_niters_no_prec;
  int expected_niters_prec;
  int niters_max_no_prec;
  int niters_max_prec;
  double normr;
};
struct TestSymmetryData {
  double depsym_spmv;
  double depsym_mg;
  int count_fail;
  double spmv_residual[2];  // |b - A xexact|_inf of the two SpMV calls
};
struct TestNormsData {
  double *values;
  double mean;
  double variance;
  int samples;
  bool pass;
};
int TestCG(SparseMatrix &A, CGData &data, Vector &b, Vector &x,
           TestCGData &testcg_data);
int TestSymmetry(SparseMatrix &A, Vector &b, Vector &xexact,
                 TestSymmetryData &testsymmetry_data);
int TestNorms(TestNormsData &testnorms_data);
int ComputeResidual(const local_int_t n, const Vector &v1, const Vector &v2,
                    double &residual);

// ---- report (ReportResults.hpp:22) --------------------------------------
struct RunSummary {
  int numberOfMgLevels;
  int numberOfCgSets;
  int refMaxIters;
  int optMaxIters;
  double refTolerance;
  double *times;  // times[0..9] as in main.cpp
  const TestCGData *testcg;
  const TestSymmetryData *testsymmetry;
  const TestNormsData *testnorms;
  int global_failure;
  bool quickPath;
  const char *ordering;
};
// Prints the run summary and writes the result file (ReportFile).
void ReportResults(const SparseMatrix &A, const RunSummary &s);
// The result file in the reference's key=value format with its section and
// key names (ReportResults.cpp:211-405, OutputFile.cpp:92-129), written as
// HPCG-Benchmark_3.1_<date>.txt in the working directory; returns the text
// and the file name through *path (empty if it could not be written).
std::string ReportFile(const SparseMatrix &A, const RunSummary &s, std::string *path);

#endif
