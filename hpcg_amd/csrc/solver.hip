// solver.hip -- host orchestration of SYMGS, the MG V-cycle and PCG over the
// device kernels.  Everything is enqueued on one stream; the host waits only
// where the reference reads a scalar (initial residual, the convergence test
// when a tolerance is set, the final residual).
#include <chrono>
#include <cmath>

#include "internal.hpp"

namespace hpcgx {

// ComputeSYMGS_ref.cpp:67-100: forward sweep, then backward sweep.
//  * lexicographic ordering: levels of the row DAG in increasing order, then
//    in decreasing order; rows of a level are independent, so this is the
//    reference's sequential sweep, row for row;
//  * multicolour ordering: colours 0..C-1, then C-1..0.
namespace {
// Event pair around a launch on the finest level (hpcgx_kernel_timing), so
// the dominant kernel's launches can be timed on their own stream.
struct LaunchLog {
  Problem &P;
  hipStream_t s;
  bool on;
  LaunchLog(Problem &P_, const Level &L, hipStream_t s_)
      : P(P_), s(s_), on(L.finest && P_.klog.on) {
    if (!on) return;
    auto &K = P.klog;
    if (K.used + 2 > K.ev.size()) {
      for (int k = 0; k < 64; ++k) {
        hipEvent_t e;
        HPCGX_HIP(hipEventCreate(&e));
        K.ev.push_back(e);
      }
    }
    HPCGX_HIP(hipEventRecord(K.ev[K.used], s));
  }
  void done() {
    if (!on) return;
    HPCGX_HIP(hipEventRecord(P.klog.ev[P.klog.used + 1], s));
    P.klog.used += 2;
  }
};

// One colour step (one launch per colour).
void colour_step(Problem &P, const Level &L, int c, const double *r, double *x,
                 hipStream_t s, double *dpart = nullptr) {
  LaunchLog log(P, L, s);
  launch_symgs_rows(P, L, L.color_ptr[c], L.color_ptr[c + 1], r, x, s, dpart);
  log.done();
}

// A whole sweep direction as one wavefront launch.
void wave_sweep(Problem &P, const Level &L, int dir, const double *r, double *x,
                hipStream_t s) {
  LaunchLog log(P, L, s);
  launch_symgs_wave(P, L, dir, r, x, s);
  log.done();
}
}  // namespace

// Whether level l's first forward sweep of a V-cycle can start from an
// unfilled x (launch_symgs_rows_zero for every colour step).
bool zero_guess_sweep(const Problem &P, int l, int ordering) {
  const Level &L = P.L[l];
  if (!P.zero_guess || ordering != HPCGX_ORDER_MULTICOLOR || !L.zmask ||
      (L.wave && P.use_wave) || !use_col16(P, L, 0))
    return false;
  for (int c = 0; c < L.ncolors; ++c)
    if (L.color_ptr[c] % kTile) return false;
  return true;
}

// The backward sweep visits every row once and leaves its final x, so with
// rz each of its colour steps adds its r.z partials: colour c's at rz + the
// partial counts of the colours visited before it.  zero_x: x was zero
// when the sweep began (only with zero_guess_sweep); the forward colour
// steps then read nothing but the x they wrote.
void symgs(Problem &P, int l, int ordering, const double *r, double *x,
           hipStream_t s, double *rz, bool zero_x) {
  const Level &L = P.L[l];
  const bool wave = ordering == HPCGX_ORDER_MULTICOLOR && L.wave && P.use_wave;
  if (rz && (ordering != HPCGX_ORDER_MULTICOLOR || wave))
    throw Error(HPCGX_EINVAL, "r.z partials need the multicolour colour steps");
  if (zero_x && !zero_guess_sweep(P, l, ordering))
    throw Error(HPCGX_EINVAL, "zero-x sweep not available on this level");
  if (wave) {
    wave_sweep(P, L, 0, r, x, s);
    wave_sweep(P, L, 1, r, x, s);
  } else if (ordering == HPCGX_ORDER_MULTICOLOR) {
    for (int c = 0; c < L.ncolors; ++c) {
      if (zero_x)
        (void)launch_symgs_rows_zero(P, L, L.color_ptr[c], L.color_ptr[c + 1], r, x, s);
      else
        colour_step(P, L, c, r, x, s);
    }
    int64_t off = 0;
    for (int c = L.ncolors - 1; c >= 0; --c) {
      colour_step(P, L, c, r, x, s, rz ? rz + off : nullptr);
      off += rz_slots(L.color_ptr[c + 1] - L.color_ptr[c]);
    }
  } else {
    for (int t = 0; t < L.nsched; ++t)
      launch_symgs_list(L, L.level_rows + L.level_ptr[t],
                        L.level_ptr[t + 1] - L.level_ptr[t], r, x, s);
    for (int t = L.nsched - 1; t >= 0; --t)
      launch_symgs_list(L, L.level_rows + L.level_ptr[t],
                        L.level_ptr[t + 1] - L.level_ptr[t], r, x, s);
  }
}

// ComputeMG_ref.cpp:39-63 with the fine-level SpMV and the restriction fused
// (only the injected rows of A*x are computed).  x = 0 (:42) is a fill
// kernel, or, where zero_guess_sweep allows, implied: the first forward
// sweep then reads only the x it has written.
void mg(Problem &P, int l, int ordering, const double *r, double *x,
        hipStream_t s, double *rz) {
  Level &L = P.L[l];
  const bool zg = zero_guess_sweep(P, l, ordering);
  if (!zg) launch_fill(x, L.n, 0.0, s);                 // :42
  if (l + 1 < P.nlevels) {
    symgs(P, l, ordering, r, x, s, nullptr, zg);        // :47
    launch_residual_restriction(P, L, r, x, L.rc, s);   // :49-51
    mg(P, l + 1, ordering, L.rc, L.xc, s);              // :52
    launch_prolongation(L, L.xc, x, s);                 // :53
    symgs(P, l, ordering, r, x, s, rz);                 // :55
  } else {
    symgs(P, l, ordering, r, x, s, rz, zg);             // :59
  }
}

int64_t rz_partials(const Problem &P, int ordering) {
  const Level &L = P.L[0];
  if (!P.rz_fused || !P.rz_part || ordering != HPCGX_ORDER_MULTICOLOR ||
      (L.wave && P.use_wave))
    return 0;
  int64_t count = 0;
  for (int c = 0; c < L.ncolors; ++c) {
    const int64_t b = L.color_ptr[c], e = L.color_ptr[c + 1];
    if (e <= b) continue;
    if (!use_col16(P, L, b)) return 0;
    count += rz_slots(e - b);
  }
  return count <= P.nrz_partials ? count : 0;
}

namespace {
double read_scalar(const Problem &P, int slot, hipStream_t s) {
  double v = 0.0;
  HPCGX_HIP(hipMemcpyAsync(&v, P.scalars + slot, sizeof v,
                           hipMemcpyDeviceToHost, s));
  HPCGX_HIP(hipStreamSynchronize(s));
  return v;
}

// HIP-event timing of CG's phases (times[1..5]): an event pair around each
// phase, recorded on the launch stream; the pairs are read once, when the
// call has finished, so timing adds no host waits between phases.  The
// events come from the problem's pool and are reused by later calls.
struct PhaseTimer {
  Problem &P;
  bool on;
  hipStream_t s;
  size_t used = 0;
  std::vector<std::pair<int, size_t>> spans;  // (times[] slot, index of the start event)
  PhaseTimer(Problem &P_, bool on_, hipStream_t s_) : P(P_), on(on_), s(s_) {}
  hipEvent_t next() {
    auto &pool = P.tlog.ev;
    if (used == pool.size()) {
      hipEvent_t e;
      HPCGX_HIP(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  void tick() {
    if (on) HPCGX_HIP(hipEventRecord(next(), s));
  }
  void tock(int slot) {
    if (!on) return;
    spans.emplace_back(slot, used - 1);
    HPCGX_HIP(hipEventRecord(next(), s));
  }
  // adds each phase's time to acc[slot] (seconds); waits for the stream
  void read(double *acc) {
    if (!on || spans.empty()) return;
    HPCGX_HIP(hipEventSynchronize(P.tlog.ev[used - 1]));
    for (const auto &sp : spans) {
      float ms = 0.f;
      HPCGX_HIP(hipEventElapsedTime(&ms, P.tlog.ev[sp.second], P.tlog.ev[sp.second + 1]));
      acc[sp.first] += ms * 1e-3;
    }
  }
};
}  // namespace

void check_status(Problem &P) {
  if (!P.wave_err) return;
  int e = 0;
  HPCGX_HIP(hipMemcpy(&e, P.wave_err, sizeof e, hipMemcpyDeviceToHost));
  if (e) {
    HPCGX_HIP(hipMemset(P.wave_err, 0, sizeof(int)));
    throw Error(HPCGX_EHIP, "wavefront sweep: a dependency wait timed out");
  }
}

// One CG iteration after the first (CG.cpp:102-124), enqueued on s.
void cg_iteration(Problem &P, int ordering, bool precond, double *x, hipStream_t s) {
  const Level &A = P.L[0];
  const int64_t n = A.n;
  const int64_t nrz = precond ? rz_partials(P, ordering) : 0;
  if (precond)
    mg(P, 0, ordering, P.r, P.z, s, nrz ? P.rz_part : nullptr);
  else
    launch_copy(P.r, P.z, n, s);
  if (nrz)
    launch_partials_dot(P, P.rz_part, nrz, kFinishBeta, s);
  else
    launch_dot(P, n, P.r, P.z, nullptr, kFinishBeta, s);
  if (P.defer_x)
    launch_update_xp(n, x, P.z, P.p, P.scalars, s);
  else
    launch_update_p(n, P.z, P.p, P.scalars, s, P.stream_p);
  launch_spmv_dot(P, A, P.p, P.Ap, s);
  launch_update_xr(P, n, P.defer_x ? nullptr : x, P.r, P.p, P.Ap, s);
}

namespace {
// The captured body of iterations 2.. (rebuilt when its key changes).
hipGraphExec_t iteration_graph(Problem &P, int ordering, bool precond, double *x) {
  auto &G = P.graph;
  if (G.exec && G.x == x && G.ordering == ordering && G.precond == precond) return G.exec;
  G.reset();
  if (!G.cap) HPCGX_HIP(hipStreamCreateWithFlags(&G.cap, hipStreamNonBlocking));
  HPCGX_HIP(hipStreamBeginCapture(G.cap, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    cg_iteration(P, ordering, precond, x, G.cap);
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
  return G.exec;
}
}  // namespace

// CG.cpp:59-143.  With the reductions fused into their producers:
//   r.z            -> partials from the V-cycle's last colour sweep, summed by
//                     a launch that also forms beta = rtz/oldrtz (:112-114);
//                     a dot kernel where no such partials are made
//   Ap = A p, p.Ap -> one SpMV kernel that also forms alpha (:118-120)
//   x, r updates, r.r -> one kernel (:121-124); with defer_x the r update
//                     and r.r, x's step (:121) being taken with the next
//                     p update (:115) in one pass over x and p, and after
//                     the last iteration on its own
// With HPCGX_CG_GRAPH (and no timers or wavefront sweeps) iterations 2..
// replay one captured graph of cg_iteration; with a tolerance the host reads
// the residual after each replay.
void cg(Problem &P, int ordering, const double *b, double *x, int max_iter,
        double tol, bool precond, int flags, int *niters, double *normr,
        double *normr0, double *times, hipStream_t s) {
  const auto t_begin = std::chrono::steady_clock::now();
  const Level &A = P.L[0];
  const int64_t n = A.n;
  double tp[6] = {0, 0, 0, 0, 0, 0};  // times[1..5] of this call
  PhaseTimer T(P, (flags & HPCGX_CG_TIMERS) != 0, s);

  *niters = 0;
  launch_copy(x, P.p, n, s);                             // :86
  T.tick();
  launch_spmv(P, A, P.p, P.Ap, s);                          // :87
  T.tock(3);
  T.tick();
  launch_residual0(P, n, b, P.Ap, P.r, s);               // :88-89
  T.tock(2);
  double nr = std::sqrt(read_scalar(P, kNormr2, s));
  *normr0 = nr;
  const bool check_each = tol > 0.0;
  bool wave = false;
  for (int l = 0; l < P.nlevels; ++l) wave = wave || (P.L[l].wave && P.use_wave);
  const bool graphed = (flags & HPCGX_CG_GRAPH) && !T.on && !P.klog.on && !wave;
  const int64_t nrz = precond ? rz_partials(P, ordering) : 0;
  const bool defer = P.defer_x;
  int k = 1;
  for (; k <= max_iter && nr / *normr0 > tol; ++k) {     // :100
    if (graphed && k >= 2) {
      HPCGX_HIP(hipGraphLaunch(iteration_graph(P, ordering, precond, x), s));
      *niters = k;
      if (check_each || k == max_iter) nr = std::sqrt(read_scalar(P, kNormr2, s));
      continue;
    }
    T.tick();
    if (precond)
      mg(P, 0, ordering, P.r, P.z, s, nrz ? P.rz_part : nullptr);  // :103
    else
      launch_copy(P.r, P.z, n, s);                       // :105
    T.tock(5);
    T.tick();
    if (nrz)                                             // :110 / :113-114
      launch_partials_dot(P, P.rz_part, nrz, kFinishBeta, s);
    else
      launch_dot(P, n, P.r, P.z, nullptr, kFinishBeta, s);
    T.tock(1);
    T.tick();
    if (k == 1)
      launch_copy(P.z, P.p, n, s);                       // :109
    else if (defer)
      launch_update_xp(n, x, P.z, P.p, P.scalars, s);    // :121 of k-1, :115
    else
      launch_update_p(n, P.z, P.p, P.scalars, s, P.stream_p);        // :115
    T.tock(2);
    T.tick();
    launch_spmv_dot(P, A, P.p, P.Ap, s);                 // :118-120
    T.tock(3);
    T.tick();
    launch_update_xr(P, n, defer ? nullptr : x, P.r, P.p, P.Ap, s);  // :121-123
    T.tock(2);
    *niters = k;
    if (check_each || k == max_iter) nr = std::sqrt(read_scalar(P, kNormr2, s));
  }
  if (defer && *niters > 0) {  // the last iteration's x += alpha p (:121)
    T.tick();
    launch_step_x(n, x, P.p, P.scalars, s);
    T.tock(2);
  }
  if (!check_each) nr = std::sqrt(read_scalar(P, kNormr2, s));
  HPCGX_HIP(hipStreamSynchronize(s));
  *normr = nr;
  T.read(tp);
  if (times) {
    times[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                              t_begin).count();
    for (int k = 1; k < 6; ++k) times[k] += tp[k];
  }
  check_status(P);
  if (!std::isfinite(nr))
    throw Error(HPCGX_EINVAL, "CG residual is not finite");
}

}  // namespace hpcgx
