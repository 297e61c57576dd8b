// surface_check.cpp -- test harness for the C++ surface (hpcg_amd/host):
// generates the problem with the driver's generator, optimizes it, applies
// each optimized function to inputs read from files and writes the outputs
// for the Python test to compare with the oracle.
//
// usage: surface_check nx ny nz dir
//   reads  dir/x.bin dir/y.bin dir/r.bin  (n doubles each, natural order)
//   writes dir/<name>.bin and dir/scalars.txt
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../hpcg_amd/host/DeviceProblem.hpp"
#include "../../hpcg_amd/host/OptimizedKernels.hpp"
#include "../../hpcg_amd/host/driver/Driver.hpp"

static std::string g_dir;

static bool load(const char *name, Vector &v) {
  FILE *f = std::fopen((g_dir + "/" + name).c_str(), "rb");
  if (!f) return false;
  size_t got = std::fread(v.values, sizeof(double), v.localLength, f);
  std::fclose(f);
  return got == (size_t)v.localLength;
}

static void save(const char *name, const Vector &v) {
  FILE *f = std::fopen((g_dir + "/" + name).c_str(), "wb");
  std::fwrite(v.values, sizeof(double), v.localLength, f);
  std::fclose(f);
}

#define CHECK(call)                                              \
  do {                                                           \
    int rc_ = (call);                                            \
    if (rc_) {                                                   \
      std::fprintf(stderr, "%s failed: %d\n", #call, rc_);       \
      return 1;                                                  \
    }                                                            \
  } while (0)

int main(int argc, char **argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s nx ny nz dir\n", argv[0]);
    return 2;
  }
  const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]);
  g_dir = argv[4];
  Geometry *geom = new Geometry;
  GenerateGeometry(1, 0, 1, 0, 0, 0, nx, ny, nz, 1, 1, 1, geom);
  SparseMatrix A;
  InitializeSparseMatrix(A, geom);
  Vector b, x0, xexact;
  GenerateProblem(A, &b, &x0, &xexact);
  SparseMatrix *cur = &A;
  for (int l = 1; l < 4; ++l) {
    GenerateCoarseProblem(*cur);
    cur = cur->Ac;
  }
  const local_int_t n = A.localNumberOfRows;
  CGData data;
  InitializeSparseCGData(A, data);
  CHECK(OptimizeProblem(A, data, b, x0, xexact));
  std::printf("optimized %d x %d x %d\n", nx, ny, nz);
  std::fflush(stdout);

  Vector x, y, r, out;
  InitializeVector(x, n);
  InitializeVector(y, n);
  InitializeVector(r, n);
  InitializeVector(out, n);
  if (!load("x.bin", x) || !load("y.bin", y) || !load("r.bin", r)) {
    std::fprintf(stderr, "missing inputs in %s\n", g_dir.c_str());
    return 2;
  }
  FILE *sc = std::fopen((g_dir + "/scalars.txt").c_str(), "w");
  bool opt = true;
  double t4 = 0;

  CHECK(ComputeSPMV(A, x, out));
  save("spmv.bin", out);
  CopyVector(x, out);
  CHECK(ComputeSYMGS(A, r, out));
  save("symgs.bin", out);
  CHECK(ComputeMG(A, r, out));
  save("mg.bin", out);
  CHECK(ComputeWAXPBY(n, 2.5, x, -0.75, y, out, opt));
  save("waxpby.bin", out);
  CHECK(ComputeWAXPBY(n, 1.0, x, 0.3, y, out, opt));
  save("waxpby_a1.bin", out);
  CopyVector(x, out);  // w aliasing y, as CG.cpp:121-122 does
  CHECK(ComputeWAXPBY(n, 0.7, y, 1.0, out, out, opt));
  save("waxpby_alias.bin", out);
  double dxy = 0, dxx = 0;
  CHECK(ComputeDotProduct(n, x, y, dxy, t4, opt));
  CHECK(ComputeDotProduct(n, x, x, dxx, t4, opt));
  std::fprintf(sc, "dot_xy %.17g\ndot_xx %.17g\n", dxy, dxx);

  // coarse level through its own SparseMatrix
  SparseMatrix &A1 = *A.Ac;
  const local_int_t n1 = A1.localNumberOfRows;
  Vector x1, out1;
  InitializeVector(x1, n1);
  InitializeVector(out1, n1);
  for (local_int_t i = 0; i < n1; ++i) x1.values[i] = x.values[i];
  CHECK(ComputeSPMV(A1, x1, out1));
  save("spmv1.bin", out1);
  CHECK(ComputeMG(A1, x1, out1));
  save("mg1.bin", out1);

  // diagonal doubled through Copy/ReplaceMatrixDiagonal, then restored
  Vector d, d2;
  InitializeVector(d, n);
  InitializeVector(d2, n);
  CopyMatrixDiagonal(A, d);
  for (local_int_t i = 0; i < n; ++i) d2.values[i] = 2.0 * d.values[i];
  ReplaceMatrixDiagonal(A, d2);
  CHECK(ComputeSPMV(A, x, out));
  save("spmv_diag2.bin", out);
  CHECK(ComputeMG(A, r, out));
  save("mg_diag2.bin", out);
  ReplaceMatrixDiagonal(A, d);
  CHECK(ComputeSPMV(A, x, out));
  save("spmv_restored.bin", out);

  // CG: 10 preconditioned iterations from zero, then to a tolerance
  int niters = 0;
  double normr = 0, normr0 = 0, times[10] = {0};
  ZeroVector(out);
  CHECK(CG(A, data, b, out, 10, 0.0, niters, normr, normr0, times, true));
  save("cg10.bin", out);
  std::fprintf(sc, "cg10 %d %.17g %.17g\n", niters, normr, normr0);
  ZeroVector(out);
  CHECK(CG(A, data, b, out, 10, 0.0, niters, normr, normr0, times, false));
  save("cg10_noprec.bin", out);
  std::fprintf(sc, "cg10_noprec %d %.17g %.17g\n", niters, normr, normr0);
  ZeroVector(out);
  CHECK(CG(A, data, b, out, 500, 1e-8, niters, normr, normr0, times, true));
  std::fprintf(sc, "cg_tol %d %.17g %.17g\n", niters, normr, normr0);
  std::fprintf(sc, "memory %.17g\n", OptimizeProblemMemoryUse(A));
  std::fclose(sc);

  for (Vector *v : {&x, &y, &r, &out, &x1, &out1, &d, &d2}) DeleteVector(*v);
  DeleteCGData(data);
  DeleteVector(b);
  DeleteVector(x0);
  DeleteVector(xexact);
  DeleteMatrix(A);
  std::printf("surface_check done\n");
  return 0;
}
