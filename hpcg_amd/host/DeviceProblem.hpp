// DeviceProblem.hpp -- the device copy of an HPCG problem behind the C++
// surface.
//
// OptimizeProblem hands the host CSR hierarchy (A, A.Ac, ...) to
// hpcgx_optimize_problem, which builds the device problem: column-major
// 27-slot ELL per level, colour and level schedules, rows reordered by
// colour.  Each level's SparseMatrix::optimizationData then points at a
// DeviceLevel.  The context also owns the scratch the per-call Compute*
// functions copy through (no allocation per call), all on the default HIP
// stream.
#ifndef HPCG_AMD_DEVICEPROBLEM_HPP
#define HPCG_AMD_DEVICEPROBLEM_HPP

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "hpcg_amd.h"  // <repo>/include, on the compiler's -I path
#include "SparseMatrix.hpp"
#include "Vector.hpp"

namespace hpcg_amd {

struct DeviceContext;

struct DeviceLevel {
  DeviceContext *ctx;
  int level;
};

struct DeviceContext {
  hpcgx_problem *problem;
  int nlevels;
  int64_t nrow[8];
  DeviceLevel levels[8];
  SparseMatrix *fine;    // the matrix whose DeleteMatrix releases this
  double *stage;         // natural-order staging, nrow[0] doubles
  double *vec[4];        // device-order scratch vectors, nrow[0] doubles each
  double build_seconds;  // wall time of hpcgx_optimize_problem + scratch
  double *pinned[2];     // pinned host staging of the vector copies (kCopyChunk doubles each)
  hipEvent_t copied[2];  // a staging buffer's last copy has finished
};

// Vector copies move in chunks of this many doubles through two pinned
// buffers: host threads fill one while the other crosses PCIe.
constexpr int64_t kCopyChunk = (int64_t)1 << 23;

// Row orderings of the Gauss-Seidel sweeps: the optimized functions use
// the multicolour schedule; the reference-ordering functions of the driver
// use the natural (lexicographic) order.
extern int optimized_ordering;

// Record HIP-event times of CG's phases (times[1..5]); on by default.  The
// events are read once per CG call, but timed iterations launch one by one
// instead of replaying a captured graph.
extern bool cg_phase_timers;
// Give OptimizeProblem each level's grid dimensions (default) and keep the
// affine offset words.  Off: the schedules come from the row-by-row
// traversal of the matrix alone and every narrow offset word is loaded, so
// nothing uses what is known of the grid beyond the matrix itself.
extern bool grid_knowledge;

// Builds the device problem for A's hierarchy unless it exists; returns
// nullptr (after printing why) on failure.
DeviceContext *EnsureDeviceProblem(SparseMatrix &A);
// The level behind a matrix, or nullptr before OptimizeProblem.
const DeviceLevel *LevelOf(const SparseMatrix &A);
// As LevelOf, but builds the device problem first when A is a fine level
// that has none (the reference calls its *_ref kernels before
// OptimizeProblem); prints why and returns nullptr on failure.
const DeviceLevel *LevelFor(const SparseMatrix &A, const char *where);

// The operations behind the optimized functions, with the Gauss-Seidel
// ordering as a parameter (HPCGX_ORDER_*).
int DeviceSYMGS(const SparseMatrix &A, const Vector &r, Vector &x, int ordering);
int DeviceMG(const SparseMatrix &A, const Vector &r, Vector &x, int ordering);
int DeviceCG(const SparseMatrix &A, const Vector &b, Vector &x, int max_iter,
             double tolerance, int &niters, double &normr, double &normr0,
             double *times, bool doPreconditioning, int ordering);
// The most recently built context (Dot and WAXPBY carry no matrix).
DeviceContext *CurrentContext();

// Host (natural order) <-> device order on a level; synchronous.
int Upload(const DeviceLevel &L, const Vector &v, double *dev);
int Download(const DeviceLevel &L, const double *dev, Vector &v);
// Plain copies of n values with no reordering; synchronous.
int UploadRaw(const double *host, double *dev, int64_t n);
int DownloadRaw(const double *dev, double *host, int64_t n);

// Prints "hpcg_amd: <where>: <last error>" and returns rc (nonzero).
int Report(const char *where, int rc);

}  // namespace hpcg_amd

#endif
