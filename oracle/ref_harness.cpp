// ref_harness.cpp -- drives the compiled HPCG 3.1 reference to produce golden
// vectors and CPU-baseline timings.  TEST INFRASTRUCTURE ONLY.
//
// It is linked against the reference's own src/*.cpp (all but main.cpp) by
// oracle/Makefile; the binaries land in oracle/_ref/ (git-ignored).  Only the
// harness itself is part of this repository; no reference source is copied.
//
//   ref_golden <nx> <ny> <nz> <outdir>   dump per-kernel vectors, CG
//                                        trajectories, TestCG/TestSymmetry
//                                        results and the flop/byte model
//   ref_time   <nx> <ny> <nz> <iters>    time one CG_ref set (main.cpp:212-224)
//                                        and print a JSON line
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <unistd.h>

#include "CG.hpp"
#include "CG_ref.hpp"
#include "CGData.hpp"
#include "ComputeDotProduct_ref.hpp"
#include "ComputeMG_ref.hpp"
#include "ComputeProlongation_ref.hpp"
#include "ComputeRestriction_ref.hpp"
#include "ComputeSPMV_ref.hpp"
#include "ComputeSYMGS_ref.hpp"
#include "ComputeWAXPBY_ref.hpp"
#include "GenerateCoarseProblem.hpp"
#include "GenerateGeometry.hpp"
#include "GenerateProblem.hpp"
#include "Geometry.hpp"
#include "MGData.hpp"
#include "ReportResults.hpp"
#include "SetupHalo.hpp"
#include "SparseMatrix.hpp"
#include "TestCG.hpp"
#include "TestNorms.hpp"
#include "TestSymmetry.hpp"
#include "Vector.hpp"
#include "mytimer.hpp"

namespace {

std::string g_dir;
std::string g_manifest;

void put(const std::string &name, const void *p, size_t bytes,
         const char *dtype, size_t count) {
  const std::string path = g_dir + "/" + name + ".bin";
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  std::fwrite(p, 1, bytes, f);
  std::fclose(f);
  char buf[256];
  std::snprintf(buf, sizeof buf, "%s\"%s\": [\"%s\", %zu]",
                g_manifest.empty() ? "" : ",\n", name.c_str(), dtype, count);
  g_manifest += buf;
}
void put_f64(const std::string &n, const double *p, size_t c) { put(n, p, c * 8, "f64", c); }
void put_i32(const std::string &n, const int32_t *p, size_t c) { put(n, p, c * 4, "i32", c); }
void put_u8(const std::string &n, const uint8_t *p, size_t c) { put(n, p, c, "u8", c); }

// splitmix64 -> uniform in [-1, 1): an exactly reproducible input stream.
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (2.0 / 9007199254740992.0) - 1.0; }
};

void fill_uniform(Vector &v, uint64_t seed) {
  SplitMix g{seed};
  for (local_int_t i = 0; i < v.localLength; ++i) v.values[i] = g.uniform();
}

// main.cpp:120-144: geometry, level 0, halo, 3 coarse levels.
void build(int nx, int ny, int nz, SparseMatrix &A, Vector &b, Vector &x,
           Vector &xexact) {
  Geometry *geom = new Geometry;
  GenerateGeometry(1, 0, 1, 0, 0, 0, nx, ny, nz, 0, 0, 0, geom);
  InitializeSparseMatrix(A, geom);
  GenerateProblem(A, &b, &x, &xexact);
  SetupHalo(A);
  SparseMatrix *cur = &A;
  for (int level = 1; level < 4; ++level) {
    GenerateCoarseProblem(*cur);
    cur = cur->Ac;
  }
}

void dump_level(const SparseMatrix &A, int level) {
  const local_int_t n = A.localNumberOfRows;
  std::vector<uint8_t> nnz(n), dslot(n);
  std::vector<int32_t> cols((size_t)n * 27);
  std::vector<double> vals((size_t)n * 27);
  for (local_int_t i = 0; i < n; ++i) {
    nnz[i] = (uint8_t)A.nonzerosInRow[i];
    dslot[i] = (uint8_t)(A.matrixDiagonal[i] - A.matrixValues[i]);
    for (int j = 0; j < 27; ++j) {
      const bool used = j < A.nonzerosInRow[i];
      cols[(size_t)i * 27 + j] = used ? A.mtxIndL[i][j] : i;
      vals[(size_t)i * 27 + j] = used ? A.matrixValues[i][j] : 0.0;
    }
  }
  const std::string p = "L" + std::to_string(level) + "_";
  put_u8(p + "nnz_row", nnz.data(), n);
  put_u8(p + "diag_slot", dslot.data(), n);
  put_i32(p + "cols", cols.data(), cols.size());
  put_f64(p + "vals", vals.data(), vals.size());
  if (A.mgData) {
    put_i32(p + "f2c", A.mgData->f2cOperator, A.mgData->rc->localLength);
  }
}

int golden(int nx, int ny, int nz, const char *outdir) {
  g_dir = outdir;
  SparseMatrix A;
  Vector b, x, xexact;
  build(nx, ny, nz, A, b, x, xexact);
  const local_int_t n = A.localNumberOfRows;

  {  // structure of every level
    const SparseMatrix *cur = &A;
    for (int l = 0; cur; ++l, cur = cur->Ac) dump_level(*cur, l);
  }
  put_f64("b", b.values, n);
  put_f64("xexact", xexact.values, n);

  Vector u1, u2, u3, y, w;
  InitializeVector(u1, n); InitializeVector(u2, n); InitializeVector(u3, n);
  InitializeVector(y, n); InitializeVector(w, n);
  fill_uniform(u1, 1); fill_uniform(u2, 2); fill_uniform(u3, 3);
  put_f64("u1", u1.values, n);
  put_f64("u2", u2.values, n);
  put_f64("u3", u3.values, n);

  // ComputeSPMV_ref
  ComputeSPMV_ref(A, u1, y);
  put_f64("spmv_u1", y.values, n);
  ComputeSPMV_ref(A, xexact, y);
  put_f64("spmv_xexact", y.values, n);

  // ComputeSYMGS_ref from a nonzero and from a zero initial guess
  CopyVector(u2, w);
  ComputeSYMGS_ref(A, u1, w);
  put_f64("symgs_u1_u2", w.values, n);
  ZeroVector(w);
  ComputeSYMGS_ref(A, u1, w);
  put_f64("symgs_u1_0", w.values, n);
  CopyVector(u3, w);
  ComputeSYMGS_ref(A, b, w);
  put_f64("symgs_b_u3", w.values, n);

  // ComputeMG_ref
  ComputeMG_ref(A, u1, w);
  put_f64("mg_u1", w.values, n);
  ComputeMG_ref(A, b, w);
  put_f64("mg_b", w.values, n);

  // Restriction (Axf = A*u2, rc = u1[f2c] - Axf[f2c]) and prolongation
  {
    MGData &mg = *A.mgData;
    ComputeSPMV_ref(A, u2, *mg.Axf);
    ComputeRestriction_ref(A, u1);
    put_f64("restrict_u1_u2", mg.rc->values, mg.rc->localLength);
    for (local_int_t i = 0; i < mg.xc->localLength; ++i) mg.xc->values[i] = u3.values[i];
    CopyVector(u2, w);
    ComputeProlongation_ref(A, w);
    put_f64("prolong_u3_u2", w.values, n);
  }

  // dot products (x!=y and x==y branches) and the three WAXPBY branches
  {
    double d[2], t = 0.0;
    ComputeDotProduct_ref(n, u1, u2, d[0], t);
    ComputeDotProduct_ref(n, u1, u1, d[1], t);
    put_f64("dot", d, 2);
    ComputeWAXPBY_ref(n, 1.0, u1, 0.7, u2, w);
    put_f64("waxpby_a1", w.values, n);
    ComputeWAXPBY_ref(n, 0.3, u1, 1.0, u2, w);
    put_f64("waxpby_b1", w.values, n);
    ComputeWAXPBY_ref(n, 0.3, u1, -0.6, u2, w);
    put_f64("waxpby_gen", w.values, n);
  }

  // CG_ref trajectories: normr/normr0 after exactly k iterations, k=1..K
  CGData data;
  InitializeSparseCGData(A, data);
  const int K = 50;
  std::vector<double> times(10, 0.0);
  for (int prec = 1; prec >= 0; --prec) {
    std::vector<double> traj(K + 1, 1.0);
    int niters = 0;
    double normr = 0, normr0 = 0;
    for (int k = 1; k <= K; ++k) {
      ZeroVector(x);
      CG_ref(A, data, b, x, k, 0.0, niters, normr, normr0, &times[0], prec == 1);
      traj[k] = normr / normr0;
    }
    put_f64(prec ? "cg_traj_prec" : "cg_traj_noprec", traj.data(), K + 1);
    if (prec) {
      put_f64("cg_x50", x.values, n);
      double nr[2] = {normr, normr0};
      put_f64("cg_normr50", nr, 2);
    }
  }

  // TestCG (TestCG.cpp:52-123), through the stub CG of CG.cpp
  {
    TestCGData tc;
    tc.count_pass = tc.count_fail = 0;
    TestCG(A, data, b, x, tc);
    int32_t out[4] = {tc.count_pass, tc.count_fail, tc.niters_max_no_prec,
                      tc.niters_max_prec};
    put_i32("testcg", out, 4);
    put_f64("testcg_normr", &tc.normr, 1);
  }

  // TestSymmetry (TestSymmetry.cpp:60-138): record the rand() vectors it
  // draws by replaying the same seed.
  {
    Vector xr, yr;
    InitializeVector(xr, n); InitializeVector(yr, n);
    std::srand(12345);
    FillRandomVector(xr);
    FillRandomVector(yr);
    put_f64("sym_x", xr.values, n);
    put_f64("sym_y", yr.values, n);
    std::srand(12345);
    TestSymmetryData ts;
    TestSymmetry(A, b, xexact, ts);
    double out[3] = {ts.depsym_spmv, ts.depsym_mg, (double)ts.count_fail};
    put_f64("testsymmetry", out, 3);
  }

  // TestNorms on a fixed sample
  {
    double v[5] = {1.0e-7, 1.1e-7, 0.9e-7, 1.05e-7, 1.0e-7};
    TestNormsData tn;
    tn.samples = 5;
    tn.values = v;
    TestNorms(tn);
    double out[3] = {tn.mean, tn.variance, tn.pass ? 1.0 : 0.0};
    put_f64("testnorms", out, 3);
  }

  // Flop/byte model: ReportResults writes HPCG-Benchmark_3.1_*.txt into cwd.
  {
    TestCGData tc{};
    TestSymmetryData ts{};
    double v[3] = {1e-7, 1e-7, 1e-7};
    TestNormsData tn;
    tn.samples = 3;
    tn.values = v;
    tn.pass = true;
    std::vector<double> t(10, 1.0);
    if (chdir(outdir) != 0) { std::perror("chdir"); return 2; }
    ReportResults(A, 4, 3, 50, 52, &t[0], tc, ts, tn, 0, true);
  }

  FILE *f = std::fopen((g_dir + "/manifest.json").c_str(), "w");
  std::fprintf(f, "{\n\"dims\": [%d, %d, %d],\n%s\n}\n", nx, ny, nz, g_manifest.c_str());
  std::fclose(f);
  return 0;
}

int time_cg(int nx, int ny, int nz, int iters) {
  SparseMatrix A;
  Vector b, x, xexact;
  const double t_setup = mytimer();
  build(nx, ny, nz, A, b, x, xexact);
  const double setup = mytimer() - t_setup;
  CGData data;
  InitializeSparseCGData(A, data);
  std::vector<double> times(10, 0.0);
  int niters = 0;
  double normr = 0, normr0 = 0;
  ZeroVector(x);
  CG_ref(A, data, b, x, iters, 0.0, niters, normr, normr0, &times[0], true);
  // HPCG flop model for one set (ReportResults.cpp:75-99)
  double nnz[4], nrow[4];
  const SparseMatrix *cur = &A;
  for (int l = 0; l < 4; ++l, cur = cur->Ac) {
    nnz[l] = (double)cur->totalNumberOfNonzeros;
    nrow[l] = (double)cur->totalNumberOfRows;
  }
  const double fit = iters, fsets = 1.0;
  double flops = (3.0 * fit + fsets) * 2.0 * nrow[0] * 2.0 + (fit + fsets) * 2.0 * nnz[0];
  for (int l = 0; l < 3; ++l) flops += fit * 10.0 * nnz[l];
  flops += fit * 4.0 * nnz[3];
  std::printf("{\"nx\": %d, \"ny\": %d, \"nz\": %d, \"iters\": %d, \"seconds\": %.6f, "
              "\"mg_seconds\": %.6f, \"spmv_seconds\": %.6f, \"setup_seconds\": %.6f, "
              "\"gflops\": %.6f, \"normr_over_normr0\": %.10e}\n",
              nx, ny, nz, niters, times[0], times[5], times[3], setup,
              flops / times[0] * 1e-9, normr / normr0);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc == 6 && std::strcmp(argv[1], "golden") == 0)
    return golden(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), argv[5]);
  if (argc == 6 && std::strcmp(argv[1], "time") == 0)
    return time_cg(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
  std::fprintf(stderr, "usage: %s golden nx ny nz outdir | time nx ny nz iters\n", argv[0]);
  return 1;
}
