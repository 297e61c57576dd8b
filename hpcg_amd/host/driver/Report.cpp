// Report.cpp -- the benchmark report (ReportResults.cpp:48-420 of the
// reference): problem summary, validation results, the operation and
// traffic counts of the reference's model and the resulting rates, as
// "Section: key = value" lines on stdout.
//
// Operation count per CG set (ReportResults.cpp:75-108), with n0, m0 the
// fine level's rows and nonzeros, m_l the nonzeros of level l, F = sets x
// iterations and S = sets:
//   DDOT 2 n0 (3F + S), WAXPBY 2 n0 (3F + S), SpMV 2 m0 (F + S),
//   MG sum over non-coarsest levels of F (4 m_l pre + 2 m_l + 4 m_l post)
//      plus 4 F m_coarsest.
// Traffic (ReportResults.cpp:110-136) counts 8-byte values and 4-byte
// indices per stored entry.  Counts are 64-bit here.
#include <cstdio>
#include <vector>

#include "../OptimizedKernels.hpp"
#include "Driver.hpp"

namespace {
struct Model {
  double ddot, waxpby, spmv, mg, total, reads, writes;
};

Model count(const std::vector<double> &nrow, const std::vector<double> &nnz, double sets,
            double iters) {
  const int L = (int)nrow.size();
  const double F = sets * iters, n0 = nrow[0], m0 = nnz[0];
  const double D = 8.0, I = 4.0;
  Model m{};
  m.ddot = (3.0 * F + sets) * 2.0 * n0;
  m.waxpby = (3.0 * F + sets) * 2.0 * n0;
  m.spmv = (F + sets) * 2.0 * m0;
  for (int l = 0; l + 1 < L; ++l) m.mg += F * 4.0 * nnz[l] + F * 2.0 * nnz[l] + F * 4.0 * nnz[l];
  m.mg += F * 4.0 * nnz[L - 1];
  m.total = m.ddot + m.waxpby + m.spmv + m.mg;

  m.reads = (3.0 * F + sets) * 2.0 * n0 * D * 2.0;  // dot and WAXPBY: two vectors in
  m.writes = (3.0 * F + sets) * D + (3.0 * F + sets) * n0 * D;
  m.reads += (F + sets) * (m0 * (D + I) + n0 * D);
  m.writes += (F + sets) * n0 * D;
  for (int l = 0; l + 1 < L; ++l) {
    const double mm = nnz[l], n = nrow[l];
    m.reads += F * (2.0 * mm * (D + I) + n * D);  // pre-smoother
    m.writes += F * n * D;
    m.reads += F * (mm * (D + I) + n * D);        // residual SpMV
    m.writes += F * mm * D;
    m.reads += F * (2.0 * mm * (D + I) + n * D);  // post-smoother
    m.writes += F * mm * D;
  }
  m.reads += F * (2.0 * nnz[L - 1] * (D + I) + nrow[L - 1] * D);
  m.writes += F * nrow[L - 1] * D;
  return m;
}

double rate(double x, double t) { return t > 0.0 ? x / t / 1.0e9 : 0.0; }
}  // namespace

void ReportResults(const SparseMatrix &A, const RunSummary &s) {
  std::vector<double> nrow, nnz;
  const SparseMatrix *M = &A;
  for (int l = 0; l < s.numberOfMgLevels && M; ++l, M = M->Ac) {
    nrow.push_back((double)M->totalNumberOfRows);
    nnz.push_back((double)M->totalNumberOfNonzeros);
  }
  const double *t = s.times;
  const double sets = s.numberOfCgSets;
  const Model m = count(nrow, nnz, sets, s.optMaxIters);
  // credit only the reference's 50 iterations per set (ReportResults.cpp:106)
  const double scale = (double)s.refMaxIters / (double)s.optMaxIters;
  const double frefnops = m.total * scale;
  const double frefbytes = (m.reads + m.writes) * scale;
  const double overhead = sets * (t[7] / 10.0 + t[9] / 10.0);
  const double totalGflops = rate(frefnops, t[0] + overhead);
  const double totalGflops24 = rate(frefnops, t[0] + sets * t[7] / 10.0);

  const TestCGData &cg = *s.testcg;
  const TestSymmetryData &sym = *s.testsymmetry;
  const TestNormsData &nrm = *s.testnorms;
  const bool valid =
      cg.count_fail == 0 && sym.count_fail == 0 && nrm.pass && !s.global_failure;

  std::printf("HPCG-AMD Benchmark: version = HPCG 3.1 semantics, MI355X device path\n");
  std::printf("Machine Summary: distributed processes = 1, threads per process = %d\n",
              A.geom->numThreads);
  std::printf("Global Problem Dimensions: nx = %lld, ny = %lld, nz = %lld\n",
              (long long)A.geom->gnx, (long long)A.geom->gny, (long long)A.geom->gnz);
  std::printf("Linear System Information: rows = %.0f, nonzeros = %.0f\n", nrow[0], nnz[0]);
  std::printf("Multigrid Information: levels = %d, pre-smoother steps = 1, "
              "post-smoother steps = 1, Gauss-Seidel ordering = %s\n",
              s.numberOfMgLevels, s.ordering);
  for (size_t l = 1; l < nrow.size(); ++l)
    std::printf("Multigrid Information: coarse grid %zu rows = %.0f, nonzeros = %.0f\n", l,
                nrow[l], nnz[l]);
  std::printf("Memory Use Information: optimized data (bytes) = %.0f\n",
              OptimizeProblemMemoryUse(A));
  std::printf("Validation Testing Phase: TestCG passed = %d, failed = %d, "
              "max iterations without / with preconditioning = %d / %d\n",
              cg.count_pass, cg.count_fail, cg.niters_max_no_prec, cg.niters_max_prec);
  std::printf("Validation Testing Phase: departure from symmetry SpMV = %.6e, MG = %.6e, "
              "failures = %d\n",
              sym.depsym_spmv, sym.depsym_mg, sym.count_fail);
  std::printf("Iteration Count Information: reference iterations per set = %d, "
              "optimized iterations per set = %d, CG sets = %d\n",
              s.refMaxIters, s.optMaxIters, s.numberOfCgSets);
  std::printf("Reproducibility Information: refTolerance = %.17g, scaled residual mean = %.6e, "
              "variance = %.6e\n",
              s.refTolerance, nrm.mean, nrm.variance);
  std::printf("Benchmark Time Summary: optimization phase = %.6f, DDOT = %.6f, WAXPBY = %.6f, "
              "SpMV = %.6f, MG = %.6f, total = %.6f (s)\n",
              t[7], t[1], t[2], t[3], t[5], t[0]);
  std::printf("Benchmark Time Summary: setup = %.6f, reference SpMV+MG = %.6f (s)\n", t[9], t[8]);
  std::printf("Floating Point Operations Summary: DDOT = %.0f, WAXPBY = %.0f, SpMV = %.0f, "
              "MG = %.0f, total = %.0f, with convergence overhead = %.0f\n",
              m.ddot, m.waxpby, m.spmv, m.mg, m.total, frefnops);
  std::printf("GB/s Summary: raw read = %.3f, raw write = %.3f, raw total = %.3f, "
              "with convergence and optimization overhead = %.3f\n",
              rate(m.reads, t[0]), rate(m.writes, t[0]), rate(m.reads + m.writes, t[0]),
              rate(frefbytes, t[0] + overhead));
  if (t[1] > 0.0)
    std::printf("GFLOP/s Summary: raw DDOT = %.3f, raw WAXPBY = %.3f, raw SpMV = %.3f, "
                "raw MG = %.3f\n",
                rate(m.ddot, t[1]), rate(m.waxpby, t[2]), rate(m.spmv, t[3]), rate(m.mg, t[5]));
  std::printf("GFLOP/s Summary: raw total = %.3f, with convergence overhead = %.3f, "
              "with convergence and optimization overhead = %.3f\n",
              rate(m.total, t[0]), rate(frefnops, t[0]), totalGflops);
  std::printf("User Optimization Overheads: optimization phase time = %.6f s, "
              "vs reference SpMV+MG time = %.3f\n",
              t[7], t[8] > 0.0 ? t[7] / t[8] : 0.0);
  if (valid) {
    std::printf("Final Summary: HPCG result is VALID with a GFLOP/s rating of = %.3f\n",
                totalGflops);
    std::printf("Final Summary: HPCG 2.4 rating for historical reasons is = %.3f\n",
                totalGflops24);
    if (s.quickPath || t[0] < 1800.0)
      std::printf("Final Summary: results are valid but execution time (s) is = %.3f "
                  "(official runs need at least 1800)\n",
                  t[0]);
  } else {
    std::printf("Final Summary: HPCG result is = INVALID\n");
  }
}
