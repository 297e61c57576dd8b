// main.cpp -- xhpcg_amd, the HPCG benchmark driver on the MI355X path.
//
// Follows the phase order of the reference's main.cpp:58-375 on one rank:
// generate the problem and 3 coarse levels, check it, time the reference
// SpMV+MG phase, run 50 reference CG iterations to fix the tolerance,
// optimize, validate (TestCG, TestSymmetry), find how many iterations the
// optimized CG needs to reach that tolerance, run the timed CG sets, check
// the norms (TestNorms) and report.
//
// The reference-ordering phases run on the device (natural Gauss-Seidel
// order), so the device problem is built before them; its build time is
// reported as part of the optimization phase (times[7]).
//
// Usage: xhpcg_amd [nx ny nz] [--nx=N --ny=N --nz=N] [--rt=seconds]
//                  [--ordering=multicolor|lexicographic] [--no-phase-timers]
//                  [--no-grid-knowledge]
// --rt=0 (the default) runs one CG set (the reference's quick path).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../DeviceProblem.hpp"
#include "../OptimizedKernels.hpp"
#include "Driver.hpp"

namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int usage(const char *argv0) {
  std::fprintf(stderr,
               "usage: %s [nx ny nz] [--nx=N] [--ny=N] [--nz=N] [--rt=SECONDS]\n"
               "          [--ordering=multicolor|lexicographic] [--no-phase-timers]\n"
               "          [--no-grid-knowledge]\n",
               argv0);
  return 2;
}
}  // namespace

int main(int argc, char *argv[]) {
  int nx = 104, ny = 104, nz = 104;
  double runningTime = 0.0;
  std::vector<int> positional;
  const char *ordering_name = "multicolor";
  for (int i = 1; i < argc; ++i) {
    const char *a = argv[i];
    if (!std::strncmp(a, "--nx=", 5)) nx = std::atoi(a + 5);
    else if (!std::strncmp(a, "--ny=", 5)) ny = std::atoi(a + 5);
    else if (!std::strncmp(a, "--nz=", 5)) nz = std::atoi(a + 5);
    else if (!std::strncmp(a, "--rt=", 5)) runningTime = std::atof(a + 5);
    else if (!std::strcmp(a, "--phase-timers")) hpcg_amd::cg_phase_timers = true;
    else if (!std::strcmp(a, "--no-phase-timers")) hpcg_amd::cg_phase_timers = false;
    else if (!std::strcmp(a, "--no-grid-knowledge")) hpcg_amd::grid_knowledge = false;
    else if (!std::strcmp(a, "--ordering=multicolor")) {
      hpcg_amd::optimized_ordering = HPCGX_ORDER_MULTICOLOR;
      ordering_name = "multicolor";
    } else if (!std::strcmp(a, "--ordering=lexicographic")) {
      hpcg_amd::optimized_ordering = HPCGX_ORDER_LEXICOGRAPHIC;
      ordering_name = "lexicographic";
    } else if (a[0] != '-') positional.push_back(std::atoi(a));
    else return usage(argv[0]);
  }
  if (positional.size() == 3) {
    nx = positional[0];
    ny = positional[1];
    nz = positional[2];
  } else if (!positional.empty()) {
    return usage(argv[0]);
  }
  const bool quickPath = runningTime == 0.0;
  const int numberOfMgLevels = 4;
  const int div = 1 << (numberOfMgLevels - 1);
  if (nx < div || ny < div || nz < div || nx % div || ny % div || nz % div) {
    std::fprintf(stderr, "nx, ny, nz must be positive multiples of %d\n", div);
    return 2;
  }
  if (CheckAspectRatio(0.125, nx, ny, nz, "local problem")) return 127;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::fprintf(stderr, "xhpcg_amd: no GPU visible; this driver has no CPU path\n");
    return 3;
  }
  std::printf("xhpcg_amd: local grid %d x %d x %d, %d MG levels, ordering %s\n", nx, ny, nz,
              numberOfMgLevels, ordering_name);
  std::fflush(stdout);

  Geometry *geom = new Geometry;
  int threads = 1;
#ifdef _OPENMP
  threads = omp_get_max_threads();
#endif
  GenerateGeometry(1, 0, threads, 0, 0, 0, nx, ny, nz, 1, 1, 1, geom);
  std::vector<double> times(10, 0.0);

  // ---- setup (main.cpp:133-146)
  double setup_time = now();
  SparseMatrix A;
  InitializeSparseMatrix(A, geom);
  Vector b, x, xexact;
  GenerateProblem(A, &b, &x, &xexact);
  SparseMatrix *cur = &A;
  for (int level = 1; level < numberOfMgLevels; ++level) {
    GenerateCoarseProblem(*cur);
    cur = cur->Ac;
  }
  times[9] = now() - setup_time;
  cur = &A;
  int ierr = CheckProblem(A, &b, &x, &xexact);
  for (cur = A.Ac; cur; cur = cur->Ac) ierr |= CheckProblem(*cur, 0, 0, 0);
  if (ierr) return 1;

  // ---- device problem (used by every phase below)
  hpcg_amd::DeviceContext *ctx = hpcg_amd::EnsureDeviceProblem(A);
  if (!ctx) return 1;
  const double device_build = ctx->build_seconds;

  CGData data;
  InitializeSparseCGData(A, data);
  const local_int_t nrow = A.localNumberOfRows, ncol = A.localNumberOfColumns;
  Vector x_overlap, b_computed;
  InitializeVector(x_overlap, ncol);
  InitializeVector(b_computed, nrow);
  FillRandomVector(x_overlap);

  // ---- reference SpMV+MG timing phase (main.cpp:176-190)
  int numberOfCalls = quickPath ? 1 : 10;
  double t_begin = now();
  for (int i = 0; i < numberOfCalls; ++i) {
    ierr = ComputeSPMV_ref(A, x_overlap, b_computed);
    if (ierr) std::printf("Error in call to SpMV: %d.\n", ierr);
    ierr = ComputeMG_ref(A, b_computed, x_overlap);
    if (ierr) std::printf("Error in call to MG: %d.\n", ierr);
  }
  times[8] = (now() - t_begin) / numberOfCalls;

  // ---- reference CG: 50 iterations fix the tolerance (main.cpp:214-238)
  int global_failure = 0, niters = 0, err_count = 0;
  double normr = 0.0, normr0 = 0.0;
  const int refMaxIters = 50;
  std::vector<double> ref_times(9, 0.0);
  ZeroVector(x);
  ierr = CG_ref(A, data, b, x, refMaxIters, 0.0, niters, normr, normr0, &ref_times[0], true);
  if (ierr) ++err_count;
  if (err_count) std::printf("%d error(s) in call(s) to reference CG.\n", err_count);
  const double refTolerance = normr / normr0;
  std::printf("Reference CG: %d iterations, scaled residual %.17g\n", niters, refTolerance);
  std::fflush(stdout);

  // ---- OptimizeProblem (main.cpp:241-245)
  double t7 = now();
  ierr = OptimizeProblem(A, data, b, x, xexact);
  t7 = now() - t7;
  times[7] = t7 + device_build;
  if (ierr) return 1;

  // ---- validation (main.cpp:262-267)
  TestCGData testcg_data{};
  TestCG(A, data, b, x, testcg_data);
  TestSymmetryData testsymmetry_data{};
  TestSymmetry(A, b, xexact, testsymmetry_data);
  std::fflush(stdout);

  // ---- optimized CG to the reference tolerance (main.cpp:277-318)
  int tolerance_failures = 0, optNiters = refMaxIters;
  const int optMaxItersCap = 10 * refMaxIters;
  double opt_worst_time = 0.0;
  std::vector<double> opt_times(9, 0.0);
  err_count = 0;
  ZeroVector(x);
  {
    const double last = opt_times[0];
    ierr = CG(A, data, b, x, optMaxItersCap, refTolerance, niters, normr, normr0, &opt_times[0],
              true);
    if (ierr) ++err_count;
    if (normr / normr0 > refTolerance) ++tolerance_failures;
    if (niters > optNiters) optNiters = niters;
    const double cur_time = opt_times[0] - last;
    if (cur_time > opt_worst_time) opt_worst_time = cur_time;
  }
  if (err_count) std::printf("%d error(s) in call(s) to optimized CG.\n", err_count);
  if (tolerance_failures) {
    global_failure = 1;
    std::printf("Failed to reduce the residual %d times.\n", tolerance_failures);
  }
  std::printf("Optimized CG: %d iterations to reach the reference tolerance\n", niters);
  const int numberOfCgSets = int(runningTime / opt_worst_time) + 1;

  // ---- timed CG sets (main.cpp:336-348)
  const int optMaxIters = optNiters;
  TestNormsData testnorms_data{};
  testnorms_data.samples = numberOfCgSets;
  testnorms_data.values = new double[numberOfCgSets];
  for (int i = 0; i < numberOfCgSets; ++i) {
    ZeroVector(x);
    ierr = CG(A, data, b, x, optMaxIters, 0.0, niters, normr, normr0, &times[0], true);
    if (ierr) std::printf("Error in call to CG: %d.\n", ierr);
    std::printf("Call [%d] Scaled Residual [%.17g]\n", i, normr / normr0);
    std::fflush(stdout);
    testnorms_data.values[i] = normr / normr0;
  }
  double residual = 0.0;
  ComputeResidual(A.localNumberOfRows, x, xexact, residual);
  std::printf("Difference between computed and exact = %g\n", residual);
  TestNorms(testnorms_data);

  RunSummary s{};
  s.numberOfMgLevels = numberOfMgLevels;
  s.numberOfCgSets = numberOfCgSets;
  s.refMaxIters = refMaxIters;
  s.optMaxIters = optMaxIters;
  s.refTolerance = refTolerance;
  s.times = &times[0];
  s.testcg = &testcg_data;
  s.testsymmetry = &testsymmetry_data;
  s.testnorms = &testnorms_data;
  s.global_failure = global_failure;
  s.quickPath = quickPath;
  s.ordering = ordering_name;
  ReportResults(A, s);

  DeleteMatrix(A);
  DeleteCGData(data);
  DeleteVector(x);
  DeleteVector(b);
  DeleteVector(xexact);
  DeleteVector(x_overlap);
  DeleteVector(b_computed);
  delete[] testnorms_data.values;
  return 0;
}
