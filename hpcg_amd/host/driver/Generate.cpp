// Generate.cpp -- the 27-point problem and its coarse levels on the host,
// in the reference's CSR layout (arrays of row pointers).
//
// Behaviour follows GenerateGeometry.cpp:49-169 (one rank),
// GenerateProblem_ref.cpp:50-219 (row order z-y-x, columns ascending,
// diagonal 26, off-diagonal -1, b = 27 - nnz, x = 0, xexact = 1),
// SetupHalo_ref.cpp (one rank: local column ids equal global ids, no halo)
// and GenerateCoarseProblem.cpp:41-111 (halved grid, f2c injects the fine
// point (2ix, 2iy, 2iz)).
#include <sys/mman.h>

#include <cstdint>
#include <cassert>
#include <cstdio>

#include "Driver.hpp"

void GenerateGeometry(int size, int rank, int numThreads, int pz, local_int_t zl,
                      local_int_t zu, local_int_t nx, local_int_t ny, local_int_t nz,
                      int npx, int npy, int npz, Geometry *geom) {
  (void)pz;
  (void)zl;
  (void)zu;
  assert(size == 1 && rank == 0);  // this driver runs one rank per GPU
  geom->size = 1;
  geom->rank = 0;
  geom->numThreads = numThreads;
  geom->nx = nx;
  geom->ny = ny;
  geom->nz = nz;
  geom->npx = npx > 0 ? npx : 1;
  geom->npy = npy > 0 ? npy : 1;
  geom->npz = npz > 0 ? npz : 1;
  geom->pz = 0;
  geom->npartz = 1;
  geom->partz_ids = new int[1]{geom->npz};
  geom->partz_nz = new local_int_t[1]{nz};
  geom->ipx = geom->ipy = geom->ipz = 0;
  geom->gnx = nx;
  geom->gny = ny;
  geom->gnz = nz;
  geom->gix0 = geom->giy0 = geom->giz0 = 0;
}

namespace {
// Large blocks are first written by the parallel loops below; backing them
// with transparent huge pages (where the kernel offers them) cuts the
// page-fault cost of that first touch.  Only whole 2 MiB pages inside the
// block are marked, so the allocator's own bookkeeping is untouched.
template <class T>
T *prefer_huge_pages(T *p, size_t count) {
  const uintptr_t huge = (uintptr_t)2 << 20;
  const uintptr_t b = ((uintptr_t)p + huge - 1) & ~(huge - 1);
  const uintptr_t e = ((uintptr_t)(p + count)) & ~(huge - 1);
  if (e > b) (void)madvise((void *)b, e - b, MADV_HUGEPAGE);
  return p;
}
}  // namespace

// Rows live in one block per array, which DeleteMatrix frees as such only
// when built with HPCG_CONTIGUOUS_ARRAYS (SparseMatrix.hpp).
#ifndef HPCG_CONTIGUOUS_ARRAYS
#error "the driver's generator allocates contiguous row arrays: build with -DHPCG_CONTIGUOUS_ARRAYS"
#endif

void GenerateProblem(SparseMatrix &A, Vector *b, Vector *x, Vector *xexact) {
  const Geometry &g = *A.geom;
  const local_int_t nx = g.nx, ny = g.ny, nz = g.nz;
  const local_int_t n = nx * ny * nz;
  const int kMax = 27;

  A.nonzerosInRow = prefer_huge_pages(new char[n], n);
  A.mtxIndG = prefer_huge_pages(new global_int_t *[n], n);
  A.mtxIndL = prefer_huge_pages(new local_int_t *[n], n);
  A.matrixValues = prefer_huge_pages(new double *[n], n);
  A.matrixDiagonal = prefer_huge_pages(new double *[n], n);
  // one block per array; row i owns entries [27 i, 27 i + 27)
  global_int_t *indG = prefer_huge_pages(new global_int_t[(size_t)kMax * n], (size_t)kMax * n);
  local_int_t *indL = prefer_huge_pages(new local_int_t[(size_t)kMax * n], (size_t)kMax * n);
  double *vals = prefer_huge_pages(new double[(size_t)kMax * n], (size_t)kMax * n);
  if (b) InitializeVector(*b, n);
  if (x) InitializeVector(*x, n);
  if (xexact) InitializeVector(*xexact, n);

  long long nnz = 0;
#pragma omp parallel for reduction(+ : nnz) schedule(static)
  for (local_int_t iz = 0; iz < nz; ++iz) {
    for (local_int_t iy = 0; iy < ny; ++iy) {
      for (local_int_t ix = 0; ix < nx; ++ix) {
        const local_int_t row = (iz * ny + iy) * nx + ix;
        global_int_t *cg = indG + (size_t)kMax * row;
        local_int_t *cl = indL + (size_t)kMax * row;
        double *v = vals + (size_t)kMax * row;
        A.mtxIndG[row] = cg;
        A.mtxIndL[row] = cl;
        A.matrixValues[row] = v;
        int m = 0;
        for (int dz = -1; dz <= 1; ++dz) {
          const local_int_t z = iz + dz;
          if (z < 0 || z >= nz) continue;
          for (int dy = -1; dy <= 1; ++dy) {
            const local_int_t y = iy + dy;
            if (y < 0 || y >= ny) continue;
            for (int dx = -1; dx <= 1; ++dx) {
              const local_int_t xx = ix + dx;
              if (xx < 0 || xx >= nx) continue;
              const local_int_t col = (z * ny + y) * nx + xx;
              cg[m] = col;
              cl[m] = col;
              if (col == row) {
                v[m] = 26.0;
                A.matrixDiagonal[row] = v + m;
              } else {
                v[m] = -1.0;
              }
              ++m;
            }
          }
        }
        A.nonzerosInRow[row] = (char)m;
        nnz += m;
        if (b) b->values[row] = 26.0 - (double)(m - 1);
        if (x) x->values[row] = 0.0;
        if (xexact) xexact->values[row] = 1.0;
      }
    }
  }
  A.title = 0;
  A.totalNumberOfRows = n;
  A.totalNumberOfNonzeros = nnz;
  A.localNumberOfRows = n;
  A.localNumberOfColumns = n;  // one rank: no halo columns
  A.localNumberOfNonzeros = (local_int_t)(nnz > 0x7fffffffLL ? 0x7fffffff : nnz);
  A.localToGlobalMap.resize(n);
  global_int_t *l2g = A.localToGlobalMap.data();
#pragma omp parallel for schedule(static)
  for (local_int_t i = 0; i < n; ++i) l2g[i] = i;
}

void GenerateCoarseProblem(const SparseMatrix &Af) {
  const Geometry &gf = *Af.geom;
  assert(gf.nx % 2 == 0 && gf.ny % 2 == 0 && gf.nz % 2 == 0);
  const local_int_t nxc = gf.nx / 2, nyc = gf.ny / 2, nzc = gf.nz / 2;
  const local_int_t nc = nxc * nyc * nzc;
  local_int_t *f2c = new local_int_t[nc];
  for (local_int_t izc = 0; izc < nzc; ++izc)
    for (local_int_t iyc = 0; iyc < nyc; ++iyc)
      for (local_int_t ixc = 0; ixc < nxc; ++ixc)
        f2c[(izc * nyc + iyc) * nxc + ixc] = ((2 * izc) * gf.ny + 2 * iyc) * gf.nx + 2 * ixc;

  Geometry *gc = new Geometry;
  GenerateGeometry(gf.size, gf.rank, gf.numThreads, gf.pz, 0, 0, nxc, nyc, nzc, gf.npx,
                   gf.npy, gf.npz, gc);
  SparseMatrix *Ac = new SparseMatrix;
  InitializeSparseMatrix(*Ac, gc);
  GenerateProblem(*Ac, 0, 0, 0);

  Vector *rc = new Vector, *xc = new Vector, *Axf = new Vector;
  InitializeVector(*rc, Ac->localNumberOfRows);
  InitializeVector(*xc, Ac->localNumberOfColumns);
  InitializeVector(*Axf, Af.localNumberOfColumns);
  Af.Ac = Ac;
  MGData *mg = new MGData;
  InitializeMGData(f2c, rc, xc, Axf, *mg);
  Af.mgData = mg;
}

// CheckProblem.cpp:79-148: every row, column, value and vector entry is what
// the generator defines.  Returns the number of mismatches.
int CheckProblem(const SparseMatrix &A, const Vector *b, const Vector *x,
                 const Vector *xexact) {
  const Geometry &g = *A.geom;
  long long bad = 0;
  for (local_int_t iz = 0; iz < g.nz; ++iz)
    for (local_int_t iy = 0; iy < g.ny; ++iy)
      for (local_int_t ix = 0; ix < g.nx; ++ix) {
        const local_int_t row = (iz * g.ny + iy) * g.nx + ix;
        int m = 0;
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              const local_int_t z = iz + dz, y = iy + dy, xx = ix + dx;
              if (z < 0 || z >= g.nz || y < 0 || y >= g.ny || xx < 0 || xx >= g.nx) continue;
              const local_int_t col = (z * g.ny + y) * g.nx + xx;
              const double want = col == row ? 26.0 : -1.0;
              if (m >= A.nonzerosInRow[row] || A.mtxIndL[row][m] != col ||
                  A.mtxIndG[row][m] != col || A.matrixValues[row][m] != want)
                ++bad;
              if (col == row && A.matrixDiagonal[row] != A.matrixValues[row] + m) ++bad;
              ++m;
            }
        if (m != A.nonzerosInRow[row]) ++bad;
        if (b && b->values[row] != 27.0 - m) ++bad;
        if (x && x->values[row] != 0.0) ++bad;
        if (xexact && xexact->values[row] != 1.0) ++bad;
      }
  if (bad) std::fprintf(stderr, "CheckProblem: %lld mismatches\n", bad);
  return bad ? 1 : 0;
}

// CheckAspectRatio.cpp: the smallest dimension must be at least
// smallest_ratio times the largest.
int CheckAspectRatio(double smallest_ratio, int x, int y, int z, const char *what) {
  const double lo = x < y ? (x < z ? x : z) : (y < z ? y : z);
  const double hi = x > y ? (x > z ? x : z) : (y > z ? y : z);
  if (lo / hi < smallest_ratio) {
    std::fprintf(stderr, "The %s (%d x %d x %d) is too stretched: ratio %g < %g\n", what, x,
                 y, z, lo / hi, smallest_ratio);
    return 127;
  }
  return 0;
}
