// fleet.hip -- many independent HPCG problems advanced by one set of launches.
//
// A fleet is the block-diagonal union of independent problems with grids of
// their own sizes.  Level l of the fleet holds every problem's level l, with
// the natural rows concatenated problem by problem, so:
//  * setup (checks, schedules, colours, device order, narrow column ids) is
//    the single-problem setup run once over the union.  No matrix entry
//    couples two problems, so each row gets the colour and schedule level it
//    has in its own problem;
//  * SpMV, the Gauss-Seidel colour steps, restriction and prolongation are
//    the single-problem launches: one launch covers all problems, and each
//    row's arithmetic is that of its own problem (bit for bit);
//  * only CG's scalars (CG.cpp:86-131) are per problem.  Each one is a
//    reduction over that problem's rows: a problem's rows in one colour are
//    contiguous in device order (colour-major, natural order inside), so the
//    fine level is cut into pieces of at most kPieceRows rows inside one
//    (problem, colour) range.  Every piece kernel writes one partial sum per
//    piece.  The last workgroup to finish sums each problem's partials in
//    index order (one wavefront per problem) and applies CG's finishing step
//    to that problem's scalars, so the per-problem dots are deterministic.
// A problem whose residual reaches the tolerance stops updating (its x, r and
// p stay as they were, its iteration count stops) while the others go on, as
// CG.cpp:100's loop test would stop it on its own.
#include <chrono>
#include <cmath>

#include "internal.hpp"

namespace hpcgx {
namespace {

enum FleetScalar { fRtz = 0, fOldRtz, fPAp, fAlpha, fBeta, fNormr2, fNormr0, kFleetScalars };
enum FleetFinish { kFResidual0 = 0, kFBeta, kFAlpha, kFNorm };

inline unsigned grid_for(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}
__device__ inline int64_t gtid() {
  return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}

struct FleetView {
  const int32_t *pieces;
  int64_t npieces;
  const int32_t *pptr;
  int count;
  double *partials;
  unsigned *ticket;
  double *S;
  int *live;
  int *niters;
  int *nlive;
  double tol;
};

FleetView view_of(Problem &P, double tol) {
  auto &F = P.fleet;
  return FleetView{F.pieces, F.npieces, F.pptr, F.count, F.partials, F.ticket,
                   F.S,      F.live,    F.niters, F.nlive, tol};
}

// fleet_blocks 0: one workgroup per piece, each reduction finished by a
// launch of its own; else a persistent grid of that many workgroups whose
// last one finishes
unsigned piece_grid(const Problem &P) {
  const int64_t g = P.fleet.grid == 0 ? P.fleet.npieces
                                      : std::min<int64_t>(P.fleet.npieces, P.fleet.grid);
  return (unsigned)std::max<int64_t>(1, g);
}

// Workgroup sum in a fixed order: 64-lane shuffle tree, then the wavefront
// sums in order.  Valid in thread 0.
__device__ inline double fblock_sum(double v) {
  __shared__ double wsum[kBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wsum[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
  }
  __syncthreads();
  return t;
}

// After every piece's partial is stored: the last workgroup (ticket) sums
// each problem's partials and applies the finishing step to its scalars.
//   kFResidual0  |r|^2 of r = b - Ax; normr0; the problem is live while
//                normr/normr0 > tol (CG.cpp:88-100)
//   kFBeta       rtz, old rtz, beta = rtz/oldrtz (CG.cpp:110-114)
//   kFAlpha      pAp, alpha = rtz/pAp (CG.cpp:119-120)
//   kFNorm       |r|^2 after the updates; one more iteration; still live?
// Frozen problems (not live) keep their scalars.  The ticket is reset for
// the next launch and nlive counts the live problems.
// The finishing step itself, by one workgroup once every partial is
// visible: the last workgroup of a piece kernel (fleet_finish) or a launch
// of its own (k_ffinish, after a piece kernel with one workgroup per piece).
__device__ void fleet_finish_sums(const FleetView &F, int finish) {
  __shared__ int live_s[kMaxFleet];
  const int lane = threadIdx.x & 63;
  for (int q = threadIdx.x >> 6; q < F.count; q += (int)(blockDim.x >> 6)) {
    double v = 0.0;
    for (int k = F.pptr[q] + lane; k < F.pptr[q + 1]; k += 64)
      v += __hip_atomic_load(&F.partials[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) {
      double *S = F.S;
      const int C = F.count;
      int live = F.live[q];
      if (finish == kFResidual0) {
        const double nr = std::sqrt(v);
        S[fNormr2 * C + q] = v;
        S[fNormr0 * C + q] = nr;
        S[fRtz * C + q] = 0.0;
        F.niters[q] = 0;
        live = nr / nr > F.tol;
      } else if (live && finish == kFBeta) {
        const double old = S[fRtz * C + q];
        S[fOldRtz * C + q] = old;
        S[fRtz * C + q] = v;
        S[fBeta * C + q] = v / old;
      } else if (live && finish == kFAlpha) {
        S[fPAp * C + q] = v;
        S[fAlpha * C + q] = S[fRtz * C + q] / v;
      } else if (live && finish == kFNorm) {
        S[fNormr2 * C + q] = v;
        F.niters[q] += 1;
        live = std::sqrt(v) / S[fNormr0 * C + q] > F.tol;
      }
      F.live[q] = live;
      live_s[q] = live;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int q = 0; q < F.count; ++q) c += live_s[q];
    *F.nlive = c;
  }
}

__device__ void fleet_finish(const FleetView &F, int finish) {
  __shared__ int is_last;
  if (threadIdx.x == 0) {
    __threadfence();
    is_last = atomicAdd(F.ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  fleet_finish_sums(F, finish);
  if (threadIdx.x == 0) *F.ticket = 0u;
}

// The finishing step in a launch of its own (one workgroup of kFinishThreads,
// a wavefront per problem in two rounds for 32 problems): the launch
// boundary orders the piece kernel's partial stores before these loads, so
// the piece kernel needs no fence or ticket.
constexpr int kFinishThreads = 1024;
__global__ __launch_bounds__(kFinishThreads) void k_ffinish(FleetView F, int finish) {
  fleet_finish_sums(F, finish);
}

// The piece kernels: each workgroup takes pieces blockIdx.x, blockIdx.x +
// gridDim.x, ...; piece k covers device rows [beg, end) of problem q.  A
// thread visits rows beg + threadIdx.x + kBlock*j, j = 0, 1, ... in order;
// the main loops load four of them before using any (four rows' loads in
// flight per wavefront; the sums are those of one row at a time).
struct Piece {
  int beg, end, q;
};
__device__ inline Piece piece_at(const FleetView &F, int64_t k) {
  return Piece{F.pieces[3 * k], F.pieces[3 * k + 1], F.pieces[3 * k + 2]};
}

// r = b - Ax (x's product is in Ap) and |r|^2 per problem (CG.cpp:88-89)
template <bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_fresidual0(FleetView F, const double *__restrict__ b,
                                                       const double *__restrict__ Ap,
                                                       double *__restrict__ r) {
  for (int64_t k = blockIdx.x; k < F.npieces; k += gridDim.x) {
    const Piece w = piece_at(F, k);
    double s = 0.0;
    int64_t i = w.beg + threadIdx.x;
    for (; i + 3 * kBlock < w.end; i += 4 * kBlock) {
      double bv[4], av[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        bv[u] = b[i + u * kBlock];
        av[u] = Ap[i + u * kBlock];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double ri = bv[u] + -1.0 * av[u];
        r[i + u * kBlock] = ri;
        s += ri * ri;
      }
    }
    for (; i < w.end; i += kBlock) {
      const double ri = b[i] + -1.0 * Ap[i];
      r[i] = ri;
      s += ri * ri;
    }
    s = fblock_sum(s);
    if (threadIdx.x == 0) F.partials[k] = s;
  }
  if (!SPLIT) fleet_finish(F, kFResidual0);
}

// x.y per live problem, finished as `finish`
template <bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_fdot(FleetView F, const double *__restrict__ x,
                                                 const double *__restrict__ y, int finish) {
  for (int64_t k = blockIdx.x; k < F.npieces; k += gridDim.x) {
    const Piece w = piece_at(F, k);
    double s = 0.0;
    if (F.live[w.q]) {
      int64_t i = w.beg + threadIdx.x;
      for (; i + 3 * kBlock < w.end; i += 4 * kBlock) {
        double a[4], c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          a[u] = x[i + u * kBlock];
          c[u] = y[i + u * kBlock];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) s += a[u] * c[u];
      }
      for (; i < w.end; i += kBlock) s += x[i] * y[i];
    }
    s = fblock_sum(s);
    if (threadIdx.x == 0) F.partials[k] = s;
  }
  if (!SPLIT) fleet_finish(F, finish);
}

// p = z (first iteration, CG.cpp:109) or p = z + beta_q p (CG.cpp:115)
__global__ __launch_bounds__(kBlock) void k_fupdate_p(FleetView F, const double *__restrict__ z,
                                                      double *__restrict__ p, int first) {
  for (int64_t k = blockIdx.x; k < F.npieces; k += gridDim.x) {
    const Piece w = piece_at(F, k);
    if (!F.live[w.q]) continue;
    const double beta = F.S[fBeta * F.count + w.q];
    for (int64_t i = w.beg + threadIdx.x; i < w.end; i += kBlock)
      p[i] = first ? z[i] : z[i] + beta * p[i];
  }
}

// x += alpha_q p; r -= alpha_q Ap; |r|^2 (CG.cpp:121-124)
template <bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_fupdate_xr(FleetView F, double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       const double *__restrict__ p,
                                                       const double *__restrict__ Ap) {
  for (int64_t k = blockIdx.x; k < F.npieces; k += gridDim.x) {
    const Piece w = piece_at(F, k);
    double s = 0.0;
    if (F.live[w.q]) {
      const double alpha = F.S[fAlpha * F.count + w.q];
      const double malpha = -alpha;
      int64_t i = w.beg + threadIdx.x;
      for (; i + 3 * kBlock < w.end; i += 4 * kBlock) {
        double xv[4], pv[4], rv[4], av[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t j = i + u * kBlock;
          xv[u] = x[j];
          pv[u] = p[j];
          rv[u] = r[j];
          av[u] = Ap[j];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t j = i + u * kBlock;
          x[j] = xv[u] + alpha * pv[u];
          const double ri = rv[u] + malpha * av[u];
          r[j] = ri;
          s += ri * ri;
        }
      }
      for (; i < w.end; i += kBlock) {
        x[i] = x[i] + alpha * p[i];
        const double ri = r[i] + malpha * Ap[i];
        r[i] = ri;
        s += ri * ri;
      }
    }
    s = fblock_sum(s);
    if (threadIdx.x == 0) F.partials[k] = s;
  }
  if (!SPLIT) fleet_finish(F, kFNorm);
}

// out[c*(Q+1) + q] = first device row of colour c whose natural row is >=
// off[q] (a colour block lists its natural rows in increasing order).
__global__ __launch_bounds__(kBlock) void k_seg_bounds(int C, int Q, const int64_t *color_ptr,
                                                       const int32_t *iperm, const int64_t *off,
                                                       int64_t *out) {
  const int64_t k = gtid();
  if (k >= (int64_t)C * (Q + 1)) return;
  const int c = (int)(k / (Q + 1)), q = (int)(k % (Q + 1));
  int64_t lo = color_ptr[c], hi = color_ptr[c + 1];
  const int64_t key = off[q];
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if ((int64_t)iperm[mid] < key) lo = mid + 1; else hi = mid;
  }
  out[k] = lo;
}

void fleet_iteration(Problem &P, const FleetView &V, int ordering, bool precond, double *x,
                     bool first, hipStream_t s) {
  const Level &A = P.L[0];
  const unsigned G = piece_grid(P);
  if (precond)
    mg(P, 0, ordering, P.r, P.z, s);                          // CG.cpp:103
  else
    launch_copy(P.r, P.z, A.n, s);                            // :105
  const bool split = P.fleet.grid == 0;
  if (split) {
    k_fdot<true><<<G, kBlock, 0, s>>>(V, P.r, P.z, kFBeta);   // :110, :113-114
    k_ffinish<<<1, kFinishThreads, 0, s>>>(V, kFBeta);
  } else {
    k_fdot<false><<<G, kBlock, 0, s>>>(V, P.r, P.z, kFBeta);
  }
  HPCGX_LAUNCHED();
  k_fupdate_p<<<G, kBlock, 0, s>>>(V, P.z, P.p, first ? 1 : 0);  // :109, :115
  HPCGX_LAUNCHED();
  launch_spmv(P, A, P.p, P.Ap, s);                            // :118
  if (split) {
    k_fdot<true><<<G, kBlock, 0, s>>>(V, P.p, P.Ap, kFAlpha); // :119-120
    k_ffinish<<<1, kFinishThreads, 0, s>>>(V, kFAlpha);
  } else {
    k_fdot<false><<<G, kBlock, 0, s>>>(V, P.p, P.Ap, kFAlpha);
  }
  HPCGX_LAUNCHED();
  if (split) {
    k_fupdate_xr<true><<<G, kBlock, 0, s>>>(V, x, P.r, P.p, P.Ap);  // :121-124
    k_ffinish<<<1, kFinishThreads, 0, s>>>(V, kFNorm);
  } else {
    k_fupdate_xr<false><<<G, kBlock, 0, s>>>(V, x, P.r, P.p, P.Ap);
  }
  HPCGX_LAUNCHED();
}

// The captured body of iterations 2.. (rebuilt when its key changes).
hipGraphExec_t fleet_graph(Problem &P, const FleetView &V, int ordering, bool precond,
                           double *x) {
  auto &F = P.fleet;
  auto &G = F.graph;
  if (G.exec && G.x == x && G.ordering == ordering && G.precond == precond &&
      F.graph_tol == V.tol)
    return G.exec;
  G.reset();
  if (!G.cap) HPCGX_HIP(hipStreamCreateWithFlags(&G.cap, hipStreamNonBlocking));
  HPCGX_HIP(hipStreamBeginCapture(G.cap, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    fleet_iteration(P, V, ordering, precond, x, false, G.cap);
  } catch (...) {
    (void)hipStreamEndCapture(G.cap, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HPCGX_HIP(hipStreamEndCapture(G.cap, &graph));
  const hipError_t e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HPCGX_HIP(e);
  G.x = x;
  G.ordering = ordering;
  G.precond = precond;
  F.graph_tol = V.tol;
  return G.exec;
}

int read_nlive(const Problem &P, hipStream_t s) {
  int v = 0;
  HPCGX_HIP(hipMemcpyAsync(&v, P.fleet.nlive, sizeof v, hipMemcpyDeviceToHost, s));
  HPCGX_HIP(hipStreamSynchronize(s));
  return v;
}

}  // namespace

void prepare_fleet(Problem &P, const std::vector<int> &dims, hipStream_t s) {
  auto &F = P.fleet;
  const int Q = (int)(dims.size() / 3);
  HPCGX_REQUIRE(Q >= 1 && Q <= kMaxFleet, "fleet size out of range");
  F.count = Q;
  F.dims = dims;
  for (int l = 0; l < P.nlevels; ++l) {
    auto &o = F.off[l];
    o.assign(Q + 1, 0);
    for (int q = 0; q < Q; ++q)
      o[q + 1] = o[q] + (int64_t)(dims[3 * q] >> l) * (dims[3 * q + 1] >> l) *
                            (dims[3 * q + 2] >> l);
    HPCGX_REQUIRE(o[Q] == P.L[l].n, "fleet rows differ from the level's rows");
  }
  // (problem, colour) ranges of the fine level's device rows
  const Level &A = P.L[0];
  const int C = A.ncolors;
  std::vector<int64_t> seg((size_t)C * (Q + 1));
  {
    TempBufD cp(sizeof(int64_t) * (C + 1)), off(sizeof(int64_t) * (Q + 1)),
        out(sizeof(int64_t) * seg.size());
    HPCGX_HIP(hipMemcpyAsync(cp.p, A.color_ptr.data(), sizeof(int64_t) * (C + 1),
                             hipMemcpyHostToDevice, s));
    HPCGX_HIP(hipMemcpyAsync(off.p, F.off[0].data(), sizeof(int64_t) * (Q + 1),
                             hipMemcpyHostToDevice, s));
    k_seg_bounds<<<grid_for((int64_t)seg.size()), kBlock, 0, s>>>(
        C, Q, static_cast<int64_t *>(cp.p), A.iperm, static_cast<int64_t *>(off.p),
        static_cast<int64_t *>(out.p));
    HPCGX_LAUNCHED();
    HPCGX_HIP(hipMemcpyAsync(seg.data(), out.p, sizeof(int64_t) * seg.size(),
                             hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
  }
  std::vector<int32_t> pieces, pptr(Q + 1, 0);
  for (int q = 0; q < Q; ++q) {
    pptr[q] = (int32_t)(pieces.size() / 3);
    for (int c = 0; c < C; ++c) {
      const int64_t b = seg[(size_t)c * (Q + 1) + q], e = seg[(size_t)c * (Q + 1) + q + 1];
      for (int64_t k = b; k < e; k += kPieceRows) {
        pieces.push_back((int32_t)k);
        pieces.push_back((int32_t)std::min<int64_t>(k + kPieceRows, e));
        pieces.push_back(q);
      }
    }
  }
  pptr[Q] = (int32_t)(pieces.size() / 3);
  // every fine row lies in exactly one piece
  int64_t covered = 0;
  for (size_t k = 0; k < pieces.size(); k += 3) covered += pieces[k + 1] - pieces[k];
  HPCGX_REQUIRE(covered == A.n, "fleet pieces do not cover the fine level");
  F.npieces = (int64_t)pieces.size() / 3;
  F.pieces = P.arena.alloc<int32_t>(pieces.size());
  F.pptr = P.arena.alloc<int32_t>(Q + 1);
  F.partials = P.arena.alloc<double>(F.npieces);
  F.ticket = P.arena.alloc<unsigned>(1);
  F.S = P.arena.alloc<double>((size_t)kFleetScalars * Q);
  F.live = P.arena.alloc<int>(Q);
  F.niters = P.arena.alloc<int>(Q);
  F.nlive = P.arena.alloc<int>(1);
  HPCGX_HIP(hipMemcpy(F.pieces, pieces.data(), sizeof(int32_t) * pieces.size(),
                      hipMemcpyHostToDevice));
  HPCGX_HIP(hipMemcpy(F.pptr, pptr.data(), sizeof(int32_t) * (Q + 1), hipMemcpyHostToDevice));
  HPCGX_HIP(hipMemset(F.ticket, 0, sizeof(unsigned)));
  HPCGX_HIP(hipMemset(F.S, 0, sizeof(double) * kFleetScalars * Q));
  HPCGX_HIP(hipMemset(F.live, 0, sizeof(int) * Q));
  HPCGX_HIP(hipMemset(F.niters, 0, sizeof(int) * Q));
  HPCGX_HIP(hipMemset(F.nlive, 0, sizeof(int)));
}

// CG.cpp:59-143 for every problem of the fleet at once: one V-cycle, one
// SpMV and four piece kernels per iteration cover all problems; problem q
// iterates while its own |r|/|r0| > tol, up to max_iter.  With
// HPCGX_CG_GRAPH iterations 2.. replay one captured graph.  niters, normr and
// normr0 receive one entry per problem.
void cg_fleet(Problem &P, int ordering, const double *b, double *x, int max_iter, double tol,
              bool precond, int flags, int *niters, double *normr, double *normr0,
              double *times, hipStream_t s) {
  auto &F = P.fleet;
  HPCGX_REQUIRE(F.count > 0, "not a fleet");
  const auto t_begin = std::chrono::steady_clock::now();
  const Level &A = P.L[0];
  const FleetView V = view_of(P, tol);
  const unsigned G = piece_grid(P);
  launch_copy(x, P.p, A.n, s);                                 // CG.cpp:86
  launch_spmv(P, A, P.p, P.Ap, s);                             // :87
  if (P.fleet.grid == 0) {
    k_fresidual0<true><<<G, kBlock, 0, s>>>(V, b, P.Ap, P.r);  // :88-89
    k_ffinish<<<1, kFinishThreads, 0, s>>>(V, kFResidual0);
  } else {
    k_fresidual0<false><<<G, kBlock, 0, s>>>(V, b, P.Ap, P.r);
  }
  HPCGX_LAUNCHED();
  int nlive = read_nlive(P, s);
  const bool graphed = (flags & HPCGX_CG_GRAPH) && !P.klog.on && !P.use_wave;
  for (int k = 1; k <= max_iter && nlive > 0; ++k) {
    if (graphed && k >= 2)
      HPCGX_HIP(hipGraphLaunch(fleet_graph(P, V, ordering, precond, x), s));
    else
      fleet_iteration(P, V, ordering, precond, x, k == 1, s);
    if (tol > 0.0 || k == max_iter) nlive = read_nlive(P, s);
  }
  const int Q = F.count;
  std::vector<double> S((size_t)kFleetScalars * Q);
  HPCGX_HIP(hipMemcpyAsync(S.data(), F.S, sizeof(double) * S.size(), hipMemcpyDeviceToHost, s));
  HPCGX_HIP(hipMemcpyAsync(niters, F.niters, sizeof(int) * Q, hipMemcpyDeviceToHost, s));
  HPCGX_HIP(hipStreamSynchronize(s));
  bool finite = true;
  for (int q = 0; q < Q; ++q) {
    normr[q] = std::sqrt(S[(size_t)fNormr2 * Q + q]);
    normr0[q] = S[(size_t)fNormr0 * Q + q];
    finite = finite && std::isfinite(normr[q]);
  }
  if (times)
    times[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin)
                    .count();
  if (!finite) throw Error(HPCGX_EINVAL, "CG residual is not finite");
}

}  // namespace hpcgx
