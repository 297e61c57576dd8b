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
#include <ctime>
#include <list>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../OptimizedKernels.hpp"
#include "../DeviceProblem.hpp"
#include "Driver.hpp"

namespace {
// The reference's result file (OutputFile.cpp:92-129): a first line with the
// benchmark name, "version=...", then one "Section::Key=value" line per
// entry, depth first, values printed with the stream defaults.  Same
// section and key names as ReportResults.cpp:211-405, so the two files diff
// line by line.
struct KeyValue {
  std::string key, value;
  std::list<std::unique_ptr<KeyValue>> kids;
  KeyValue &add(const std::string &k, const std::string &v = "") {
    kids.emplace_back(new KeyValue{k, v, {}});
    return *kids.back();
  }
  template <class T>
  KeyValue &add(const std::string &k, T v) {
    std::ostringstream ss;
    ss << v;
    return add(k, ss.str());
  }
  KeyValue &get(const std::string &k) {  // first child with this key
    for (auto &c : kids)
      if (c->key == k) return *c;
    return add(k);
  }
  void emit(std::string &out, const std::string &prefix) const {
    out += prefix + key + "=" + value + "\n";
    for (const auto &c : kids) c->emit(out, prefix + key + "::");
  }
};

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

// Memory model of ReportResults.cpp:140-205 (reference data structures) plus
// the bytes the device problem holds.
void memory_model(const SparseMatrix &A, int nlevels, double sets, std::vector<double> &per_level,
                  double &total, double &optimized) {
  const double nnz_per_row = 27.0, fnrow = (double)A.totalNumberOfRows;
  double b = sizeof(Geometry) + sizeof(double) * sets;
  b += fnrow * (sizeof(char) + sizeof(global_int_t *) + sizeof(local_int_t *) +
                2 * sizeof(double *));
  b += fnrow * nnz_per_row * (sizeof(local_int_t) + sizeof(double) + sizeof(global_int_t));
  b += fnrow * 3 * sizeof(double);  // x, b, xexact
  const double fncol = (double)A.localNumberOfColumns;
  b += fnrow * 2 * sizeof(double) + fncol * 2 * sizeof(double);  // CGData
  per_level.assign(1, b);
  optimized = OptimizeProblemMemoryUse(A);
  total = b + optimized;
  const SparseMatrix *Af = A.Ac;
  for (int l = 1; l < nlevels && Af; ++l, Af = Af->Ac) {
    const double r = (double)Af->totalNumberOfRows, c = (double)Af->localNumberOfColumns;
    double f = r * sizeof(local_int_t) + r * sizeof(double) + 2.0 * c * sizeof(double) +
               (sizeof(Geometry) + sizeof(SparseMatrix) + 3 * sizeof(Vector) + sizeof(MGData));
    f += r * (sizeof(char) + sizeof(global_int_t *) + sizeof(local_int_t *) +
              2 * sizeof(double *));
    f += r * nnz_per_row * (sizeof(local_int_t) + sizeof(double) + sizeof(global_int_t));
    per_level.push_back(f);
    total += f;
  }
}
}  // namespace

std::string ReportFile(const SparseMatrix &A, const RunSummary &s, std::string *path) {
  std::vector<double> nrow, nnz;
  const SparseMatrix *M = &A;
  for (int l = 0; l < s.numberOfMgLevels && M; ++l, M = M->Ac) {
    nrow.push_back((double)M->totalNumberOfRows);
    nnz.push_back((double)M->totalNumberOfNonzeros);
  }
  const double *t = s.times;
  const double sets = s.numberOfCgSets;
  const Model m = count(nrow, nnz, sets, s.optMaxIters);
  const double scale = (double)s.refMaxIters / (double)s.optMaxIters;
  const double frefnops = m.total * scale;
  const double frefbytes = (m.reads + m.writes) * scale;
  const double overhead = sets * (t[7] / 10.0 + t[9] / 10.0);
  const TestCGData &cg = *s.testcg;
  const TestSymmetryData &sym = *s.testsymmetry;
  const TestNormsData &nrm = *s.testnorms;
  const bool valid = cg.count_fail == 0 && sym.count_fail == 0 && nrm.pass && !s.global_failure;
  const Geometry &g = *A.geom;

  KeyValue doc;
  doc.add("Implementation",
          std::string("hpcg_amd: HIP kernels for MI355X (gfx950), one rank per GPU") +
              (hpcg_amd::grid_knowledge ? ""
                                        : "; schedules and column offsets from the matrix "
                                          "alone (--no-grid-knowledge)"));
  doc.add("Machine Summary").add("Distributed Processes", g.size);
  doc.get("Machine Summary").add("Threads per processes", g.numThreads);
  KeyValue &gp = doc.add("Global Problem Dimensions");
  gp.add("Global nx", g.gnx);
  gp.add("Global ny", g.gny);
  gp.add("Global nz", g.gnz);
  KeyValue &pd = doc.add("Processor Dimensions");
  pd.add("npx", g.npx);
  pd.add("npy", g.npy);
  pd.add("npz", g.npz);
  KeyValue &ld = doc.add("Local Domain Dimensions");
  ld.add("nx", g.nx);
  ld.add("ny", g.ny);
  int lower = 0;
  for (int i = 0; i < g.npartz; ++i) {
    ld.add("Lower ipz", lower);
    ld.add("Upper ipz", g.partz_ids[i] - 1);
    ld.add("nz", g.partz_nz[i]);
    lower = g.partz_ids[i];
  }
  doc.add("########## Problem Summary  ##########");
  doc.add("Setup Information").add("Setup Time", t[9]);
  KeyValue &ls = doc.add("Linear System Information");
  ls.add("Number of Equations", (long long)A.totalNumberOfRows);
  ls.add("Number of Nonzero Terms", (long long)A.totalNumberOfNonzeros);
  KeyValue &mi = doc.add("Multigrid Information");
  mi.add("Number of coarse grid levels", s.numberOfMgLevels - 1);
  KeyValue &cgr = mi.add("Coarse Grids");
  const SparseMatrix *Af = &A;
  for (int l = 1; l < s.numberOfMgLevels && Af->Ac; ++l, Af = Af->Ac) {
    cgr.add("Grid Level", l);
    cgr.add("Number of Equations", (long long)Af->Ac->totalNumberOfRows);
    cgr.add("Number of Nonzero Terms", (long long)Af->Ac->totalNumberOfNonzeros);
    cgr.add("Number of Presmoother Steps", Af->mgData->numberOfPresmootherSteps);
    cgr.add("Number of Postsmoother Steps", Af->mgData->numberOfPostsmootherSteps);
  }
  doc.add("########## Memory Use Summary  ##########");
  std::vector<double> per_level;
  double mem_total = 0, mem_opt = 0;
  memory_model(A, s.numberOfMgLevels, sets, per_level, mem_total, mem_opt);
  KeyValue &mu = doc.add("Memory Use Information");
  mu.add("Total memory used for data (Gbytes)", mem_total / 1e9);
  mu.add("Memory used for OptimizeProblem data (Gbytes)", mem_opt / 1e9);
  mu.add("Bytes per equation (Total memory / Number of Equations)", mem_total / nrow[0]);
  mu.add("Memory used for linear system and CG (Gbytes)", per_level[0] / 1e9);
  KeyValue &mcg = mu.add("Coarse Grids");
  for (size_t l = 1; l < per_level.size(); ++l) {
    mcg.add("Grid Level", (int)l);
    mcg.add("Memory used", per_level[l] / 1e9);
  }
  doc.add("########## V&V Testing Summary  ##########");
  KeyValue &sc = doc.add("Spectral Convergence Tests");
  sc.add("Result", cg.count_fail == 0 ? "PASSED" : "FAILED");
  KeyValue &un = sc.add("Unpreconditioned");
  un.add("Maximum iteration count", cg.niters_max_no_prec);
  un.add("Expected iteration count", cg.expected_niters_no_prec);
  KeyValue &pr = sc.add("Preconditioned");
  pr.add("Maximum iteration count", cg.niters_max_prec);
  pr.add("Expected iteration count", cg.expected_niters_prec);
  KeyValue &ds = doc.add("Departure from Symmetry |x'Ay-y'Ax|/(2*||x||*||A||*||y||)/epsilon");
  ds.add("Result", sym.count_fail == 0 ? "PASSED" : "FAILED");
  ds.add("Departure for SpMV", sym.depsym_spmv);
  ds.add("Departure for MG", sym.depsym_mg);
  doc.add("########## Iterations Summary  ##########");
  KeyValue &ic = doc.add("Iteration Count Information");
  ic.add("Result", !s.global_failure ? "PASSED" : "FAILED");
  ic.add("Reference CG iterations per set", s.refMaxIters);
  ic.add("Optimized CG iterations per set", s.optMaxIters);
  ic.add("Total number of reference iterations", s.refMaxIters * s.numberOfCgSets);
  ic.add("Total number of optimized iterations", s.optMaxIters * s.numberOfCgSets);
  doc.add("########## Reproducibility Summary  ##########");
  KeyValue &ri = doc.add("Reproducibility Information");
  ri.add("Result", nrm.pass ? "PASSED" : "FAILED");
  ri.add("Scaled residual mean", nrm.mean);
  ri.add("Scaled residual variance", nrm.variance);
  doc.add("########## Performance Summary (times in sec) ##########");
  KeyValue &bt = doc.add("Benchmark Time Summary");
  bt.add("Optimization phase", t[7]);
  bt.add("DDOT", t[1]);
  bt.add("WAXPBY", t[2]);
  bt.add("SpMV", t[3]);
  bt.add("MG", t[5]);
  bt.add("Total", t[0]);
  KeyValue &fo = doc.add("Floating Point Operations Summary");
  fo.add("Raw DDOT", m.ddot);
  fo.add("Raw WAXPBY", m.waxpby);
  fo.add("Raw SpMV", m.spmv);
  fo.add("Raw MG", m.mg);
  fo.add("Total", m.total);
  fo.add("Total with convergence overhead", frefnops);
  KeyValue &gb = doc.add("GB/s Summary");
  gb.add("Raw Read B/W", rate(m.reads, t[0]));
  gb.add("Raw Write B/W", rate(m.writes, t[0]));
  gb.add("Raw Total B/W", rate(m.reads + m.writes, t[0]));
  gb.add("Total with convergence and optimization phase overhead", rate(frefbytes, t[0] + overhead));
  KeyValue &gf = doc.add("GFLOP/s Summary");
  gf.add("Raw DDOT", rate(m.ddot, t[1]));
  gf.add("Raw WAXPBY", rate(m.waxpby, t[2]));
  gf.add("Raw SpMV", rate(m.spmv, t[3]));
  gf.add("Raw MG", rate(m.mg, t[5]));
  gf.add("Raw Total", rate(m.total, t[0]));
  gf.add("Total with convergence overhead", rate(frefnops, t[0]));
  const double totalGflops = rate(frefnops, t[0] + overhead);
  const double totalGflops24 = rate(frefnops, t[0] + sets * t[7] / 10.0);
  gf.add("Total with convergence and optimization phase overhead", totalGflops);
  KeyValue &uo = doc.add("User Optimization Overheads");
  uo.add("Optimization phase time (sec)", t[7]);
  uo.add("Optimization phase time vs reference SpMV+MG time", t[8] > 0 ? t[7] / t[8] : 0.0);
  KeyValue &fs = doc.add("Final Summary");
  if (valid) {
    fs.add("HPCG result is VALID with a GFLOP/s rating of", totalGflops);
    fs.add("HPCG 2.4 rating for historical reasons is", totalGflops24);
    if (t[0] >= 1800.0) {
      fs.add("Please upload results from the YAML file contents to", "http://hpcg-benchmark.org");
    } else {
      fs.add("Results are valid but execution time (sec) is", t[0]);
      if (s.quickPath) {
        fs.add("You have selected the QuickPath option",
               "Results are official for legacy installed systems with confirmation from the "
               "HPCG Benchmark leaders.");
        fs.add("After confirmation please upload results from the YAML file contents to",
               "http://hpcg-benchmark.org");
      } else {
        fs.add("Official results execution time (sec) must be at least", 1800.0);
      }
    }
  } else {
    fs.add("HPCG result is", "INVALID.");
    fs.add("Please review the YAML file contents", "You may NOT submit these results for "
                                                   "consideration.");
  }

  std::string text = "HPCG-Benchmark\nversion=3.1\n";
  for (const auto &c : doc.kids) c->emit(text, "");
  std::time_t now = std::time(nullptr);
  char date[32];
  std::strftime(date, sizeof date, "%Y-%m-%d_%H-%M-%S", std::localtime(&now));
  const std::string name = std::string("HPCG-Benchmark_3.1_") + date + ".txt";
  if (FILE *f = std::fopen(name.c_str(), "w")) {
    std::fputs(text.c_str(), f);
    std::fclose(f);
    if (path) *path = name;
  }
  return text;
}

void ReportResults(const SparseMatrix &A, const RunSummary &s) {
  std::string file;
  ReportFile(A, s, &file);
  if (!file.empty()) std::printf("Result file: %s\n", file.c_str());
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
