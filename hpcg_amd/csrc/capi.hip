// capi.hip -- the extern "C" entry points of libhpcg_amd.so (include/hpcg_amd.h).
// Exceptions never cross this boundary: each entry point maps them to an
// HPCGX_E* code and records the message for hpcgx_last_error().
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include <chrono>
#include <cstdio>

#include "internal.hpp"

struct hpcgx_problem {
  hpcgx::Problem P;
};

namespace {
thread_local std::string g_err;
}  // namespace

std::string &hpcgx::last_error_slot() { return g_err; }

namespace {
using hpcgx::guard;

inline hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }

using hpcgx::Error;
using hpcgx::kSlots;

hpcgx::Level &level_of(hpcgx_problem *p, int l) {
  if (!p) throw Error(HPCGX_EINVAL, "NULL problem");
  if (l < 0 || l >= p->P.nlevels) throw Error(HPCGX_EINVAL, "level out of range");
  return p->P.L[l];
}
const hpcgx::Level &level_of(const hpcgx_problem *p, int l) {
  return level_of(const_cast<hpcgx_problem *>(p), l);
}
void check_order(int o) {
  if (o != HPCGX_ORDER_LEXICOGRAPHIC && o != HPCGX_ORDER_MULTICOLOR)
    throw Error(HPCGX_EINVAL, "unknown ordering");
}

struct DevBuf {  // device buffer freed on scope exit
  void *p = nullptr;
  void alloc(size_t bytes) { HPCGX_HIP(hipMalloc(&p, bytes ? bytes : 1)); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  ~DevBuf() { reset(); }
};

struct NaturalOwner {  // natural-order level plus the buffers behind it
  hpcgx::NaturalLevel N;
  DevBuf col, val, nnz, f2c;
  void alloc(int64_t n, int64_t nc) {
    col.alloc(sizeof(int32_t) * kSlots * n);
    val.alloc(sizeof(double) * kSlots * n);
    nnz.alloc(n);
    N.col = static_cast<int32_t *>(col.p);
    N.val = static_cast<double *>(val.p);
    N.nnz_row = static_cast<uint8_t *>(nnz.p);
    N.n = n;
    N.nc = nc;
    if (nc > 0) {
      f2c.alloc(sizeof(int32_t) * nc);
      N.f2c = static_cast<int32_t *>(f2c.p);
    }
  }
  void drop_matrix() {  // keep only f2c (needed to link levels)
    col.reset();
    val.reset();
    nnz.reset();
    N.col = nullptr;
    N.val = nullptr;
    N.nnz_row = nullptr;
  }
};

void alloc_solver_state(hpcgx::Problem &P) {
  const int64_t n = P.L[0].n;
  P.r = P.arena.alloc<double>(n);
  P.z = P.arena.alloc<double>(n);
  P.p = P.arena.alloc<double>(n);
  P.Ap = P.arena.alloc<double>(n);
  P.partials = P.arena.alloc<double>(hpcgx::kMaxPartials);
  P.ticket = P.arena.alloc<unsigned>(1);
  P.scalars = P.arena.alloc<double>(hpcgx::kNumScalars);
  P.wave_err = P.arena.alloc<int>(1);
  HPCGX_HIP(hipMemset(P.ticket, 0, sizeof(unsigned)));
  HPCGX_HIP(hipMemset(P.scalars, 0, sizeof(double) * hpcgx::kNumScalars));
  HPCGX_HIP(hipMemset(P.wave_err, 0, sizeof(int)));
  P.wave_grid = hpcgx::wave_grid_size();
  P.nchunk_partials = P.L[0].nchunks;
  P.chunk_partials = P.arena.alloc<double>(P.nchunk_partials);
  // one partial per wavefront of each finest colour step's workgroups: at
  // most one workgroup per colour more than n / kBlock (up to 64 colours)
  P.nrz_partials = (n + hpcgx::kTile - 1) / hpcgx::kTile + 64;  // one per wavefront of rows
  P.rz_part = P.arena.alloc<double>(P.nrz_partials);
  P.dot_parts = P.arena.alloc<double>(hpcgx::kDotParts);
}

// shared tail of both setup paths: schedules, reordering, links
void build_levels(hpcgx::Problem &P, std::vector<NaturalOwner> &nat, int flags,
                  hipStream_t s) {
  for (int l = 0; l < P.nlevels; ++l) {
    hpcgx::finalize_level(P, l, nat[l].N, flags, s);
    nat[l].drop_matrix();
  }
  for (int l = 0; l + 1 < P.nlevels; ++l) {
    if (nat[l].N.nc != P.L[l + 1].n)
      throw Error(HPCGX_EINVAL, "f2c length differs from the next level's rows");
    hpcgx::link_levels(P, l, nat[l].N, s);
  }
  HPCGX_HIP(hipStreamSynchronize(s));
  alloc_solver_state(P);
}

}  // namespace

extern "C" {

const char *hpcgx_last_error(void) { return g_err.c_str(); }
int hpcgx_abi_version(void) { return HPCGX_ABI_VERSION; }

int hpcgx_generate_problem(int nx, int ny, int nz, int nlevels, int flags,
                           hpcgx_problem **out) {
  return guard([&] {
    if (!out) throw Error(HPCGX_EINVAL, "NULL out");
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1) throw Error(HPCGX_EINVAL, "grid dimensions must be >= 1");
    if (nlevels < 1 || nlevels > hpcgx::kMaxLevels) throw Error(HPCGX_EINVAL, "nlevels out of range");
    const int div = 1 << (nlevels - 1);
    if (nx % div || ny % div || nz % div)  // GenerateCoarseProblem.cpp:50
      throw Error(HPCGX_EINVAL, "grid dimensions must be divisible by 2^(nlevels-1)");
    if ((int64_t)nx * ny * nz >= ((int64_t)1 << 31))
      throw Error(HPCGX_EINVAL, "rows must be < 2^31 (int32 row ids)");
    auto h = std::make_unique<hpcgx_problem>();
    hpcgx::Problem &P = h->P;
    HPCGX_HIP(hipGetDevice(&P.device));
    P.nlevels = nlevels;
    P.generated = true;
    P.L[0].finest = true;
    std::vector<NaturalOwner> nat(nlevels);
    int x = nx, y = ny, z = nz;
    hipStream_t s = nullptr;
    for (int l = 0; l < nlevels; ++l) {
      const int64_t n = (int64_t)x * y * z;
      const int64_t nc = (l + 1 < nlevels) ? n / 8 : 0;
      nat[l].alloc(n, nc);
      nat[l].N.nx = x; nat[l].N.ny = y; nat[l].N.nz = z;
      nat[l].N.parts = {hpcgx::GridPart{0, x, y, z}};
      hpcgx::generate_natural(nat[l].N, s);
      // schedules and reordering of this level before generating the next
      // keep the peak footprint at about two copies of one level
      hpcgx::finalize_level(P, l, nat[l].N, flags, s);
      nat[l].drop_matrix();
      x /= 2; y /= 2; z /= 2;
    }
    for (int l = 0; l + 1 < nlevels; ++l) hpcgx::link_levels(P, l, nat[l].N, s);
    HPCGX_HIP(hipStreamSynchronize(s));
    nat.clear();
    hpcgx::check_problem(P, nullptr, nullptr, nullptr, s);  // main.cpp:150-159
    alloc_solver_state(P);
    *out = h.release();
  });
}

int hpcgx_optimize_problem(const hpcgx_host_level *levels, int nlevels,
                           int flags, hpcgx_problem **out) {
  return guard([&] {
    if (!out || !levels) throw Error(HPCGX_EINVAL, "NULL argument");
    *out = nullptr;
    if (nlevels < 1 || nlevels > hpcgx::kMaxLevels) throw Error(HPCGX_EINVAL, "nlevels out of range");
    auto h = std::make_unique<hpcgx_problem>();
    hpcgx::Problem &P = h->P;
    HPCGX_HIP(hipGetDevice(&P.device));
    P.nlevels = nlevels;
    P.L[0].finest = true;
    std::vector<NaturalOwner> nat(nlevels);
    hipStream_t s = nullptr;
    for (int l = 0; l < nlevels; ++l) {
      const hpcgx_host_level &H = levels[l];
      const int64_t n = H.nrow;
      if (n < 1 || n >= ((int64_t)1 << 31)) throw Error(HPCGX_EINVAL, "nrow out of range");
      if (!H.nonzerosInRow || !H.mtxIndL || !H.matrixValues)
        throw Error(HPCGX_EINVAL, "level arrays missing");
      const bool has_next = l + 1 < nlevels;
      if (has_next && !H.f2cOperator) throw Error(HPCGX_EINVAL, "f2cOperator missing");
      const int64_t nc = has_next ? levels[l + 1].nrow : 0;
      nat[l].alloc(n, nc);
      nat[l].N.nx = H.nx; nat[l].N.ny = H.ny; nat[l].N.nz = H.nz;
      if (H.nx > 0 && H.ny > 0 && H.nz > 0 && (int64_t)H.nx * H.ny * H.nz == n)
        nat[l].N.parts = {hpcgx::GridPart{0, H.nx, H.ny, H.nz}};
      // column-major ELL on the device, slots in the reference's row order
      hpcgx::upload_host_level(H, nat[l].N, s);
      if (has_next) {
        for (int64_t c = 0; c < nc; ++c)
          if (H.f2cOperator[c] < 0 || H.f2cOperator[c] >= n)
            throw Error(HPCGX_EINVAL, "f2cOperator entry out of range");
        HPCGX_HIP(hipMemcpy(nat[l].N.f2c, H.f2cOperator, sizeof(int32_t) * nc, hipMemcpyHostToDevice));
      }
    }
    build_levels(P, nat, flags, s);
    *out = h.release();
  });
}

int hpcgx_problem_destroy(hpcgx_problem *p) {
  return guard([&] { delete p; });
}

int hpcgx_problem_nlevels(const hpcgx_problem *p) { return p ? p->P.nlevels : 0; }

int hpcgx_problem_info(const hpcgx_problem *p, int level, hpcgx_level_info *info) {
  return guard([&] {
    const hpcgx::Level &L = level_of(p, level);
    if (!info) throw Error(HPCGX_EINVAL, "NULL info");
    info->nx = L.nx; info->ny = L.ny; info->nz = L.nz;
    info->nrow = L.n;
    info->nnz = L.nnz;
    info->ncolors = L.ncolors;
    info->nlevels_sched = L.nsched;
    info->ncoarse = L.nc;
    info->wave = (L.wave && p->P.use_wave) ? 1 : 0;
    info->nslabs = L.nslabs;
    info->col16 = L.c16 ? 1 : 0;
    info->wide_tiles = L.wide_tiles;
    info->byte_tiles = L.byte_tiles;
  });
}

double hpcgx_optimize_problem_memory_use(const hpcgx_problem *p) {
  return p ? p->P.arena.bytes : 0.0;
}

int hpcgx_problem_vectors(hpcgx_problem *p, double *b, double *x, double *xexact,
                          void *stream) {
  return guard([&] {
    if (!p) throw Error(HPCGX_EINVAL, "NULL problem");
    if (!p->P.generated && b)
      throw Error(HPCGX_EINVAL, "b is known only for generated problems");
    hpcgx::problem_vectors(p->P, b, x, xexact, S(stream));
  });
}

int hpcgx_get_schedule(const hpcgx_problem *p, int level, int which, int32_t *out_host) {
  return guard([&] {
    const hpcgx::Level &L = level_of(p, level);
    const int32_t *src = which == 0 ? L.level_of : which == 1 ? L.color_of : nullptr;
    if (which != 0 && which != 1) throw Error(HPCGX_EINVAL, "which must be 0 or 1");
    if (!src) throw Error(HPCGX_EINVAL, "schedule ids not kept (HPCGX_KEEP_SCHEDULE)");
    if (!out_host) throw Error(HPCGX_EINVAL, "NULL output");
    HPCGX_HIP(hipMemcpy(out_host, src, sizeof(int32_t) * L.n, hipMemcpyDeviceToHost));
  });
}

int hpcgx_to_device_order(const hpcgx_problem *p, int level, const double *natural,
                          double *devorder, void *stream) {
  return guard([&] {
    const hpcgx::Level &L = level_of(p, level);
    hpcgx::launch_gather(L.iperm, natural, devorder, L.n, S(stream));
  });
}

int hpcgx_to_natural_order(const hpcgx_problem *p, int level, const double *devorder,
                           double *natural, void *stream) {
  return guard([&] {
    const hpcgx::Level &L = level_of(p, level);
    hpcgx::launch_gather(L.perm, devorder, natural, L.n, S(stream));
  });
}

int hpcgx_replace_diagonal(hpcgx_problem *p, int level, const double *diag, void *stream) {
  return guard([&] {
    hpcgx::Level &L = level_of(p, level);
    hpcgx::launch_replace_diag(L, diag, S(stream));
    hpcgx::replace_diag32(L, diag, S(stream));
  });
}

int hpcgx_copy_diagonal(const hpcgx_problem *p, int level, double *diag, void *stream) {
  return guard([&] {
    const hpcgx::Level &L = level_of(p, level);
    HPCGX_HIP(hipMemcpyAsync(diag, L.diag, sizeof(double) * L.n,
                             hipMemcpyDeviceToDevice, S(stream)));
  });
}

int hpcgx_spmv(const hpcgx_problem *p, int level, const double *x, double *y, void *stream) {
  return guard([&] { hpcgx::launch_spmv(p->P, level_of(p, level), x, y, S(stream)); });
}

int hpcgx_symgs(const hpcgx_problem *p, int level, int ordering, const double *r,
                double *x, void *stream) {
  return guard([&] {
    level_of(p, level);
    check_order(ordering);
    hpcgx::symgs(const_cast<hpcgx_problem *>(p)->P, level, ordering, r, x, S(stream));
  });
}

int hpcgx_mg(hpcgx_problem *p, int level, int ordering, const double *r, double *x,
             void *stream) {
  return guard([&] {
    level_of(p, level);
    check_order(ordering);
    hpcgx::mg(p->P, level, ordering, r, x, S(stream));
  });
}

int hpcgx_restriction(const hpcgx_problem *p, int level, const double *rf,
                      const double *Axf, double *rc, void *stream) {
  return guard([&] {
    const hpcgx::Level &L = level_of(p, level);
    if (L.nc == 0) throw Error(HPCGX_EINVAL, "no coarser level");
    hpcgx::launch_restriction(L, rf, Axf, rc, S(stream));
  });
}

int hpcgx_residual_restriction(const hpcgx_problem *p, int level, const double *rf,
                               const double *xf, double *rc, void *stream) {
  return guard([&] {
    const hpcgx::Level &L = level_of(p, level);
    if (L.nc == 0) throw Error(HPCGX_EINVAL, "no coarser level");
    hpcgx::launch_residual_restriction(p->P, L, rf, xf, rc, S(stream));
  });
}

int hpcgx_prolongation(const hpcgx_problem *p, int level, const double *xc, double *xf,
                       void *stream) {
  return guard([&] {
    const hpcgx::Level &L = level_of(p, level);
    if (L.nc == 0) throw Error(HPCGX_EINVAL, "no coarser level");
    hpcgx::launch_prolongation(L, xc, xf, S(stream));
  });
}

int hpcgx_dot(hpcgx_problem *p, int64_t n, const double *x, const double *y,
              double *result_dev, void *stream) {
  return guard([&] {
    if (!p || !result_dev) throw Error(HPCGX_EINVAL, "NULL argument");
    if (n < 0) throw Error(HPCGX_EINVAL, "negative length");
    hpcgx::launch_dot(p->P, n, x, y, result_dev, hpcgx::kFinishStore, S(stream));
  });
}

int hpcgx_dot_host(hpcgx_problem *p, int64_t n, const double *x, const double *y,
                   double *result, void *stream) {
  return guard([&] {
    if (!p || !result) throw Error(HPCGX_EINVAL, "NULL argument");
    if (n < 0) throw Error(HPCGX_EINVAL, "negative length");
    double *slot = p->P.scalars + hpcgx::kDot;
    hpcgx::launch_dot(p->P, n, x, y, slot, hpcgx::kFinishStore, S(stream));
    HPCGX_HIP(hipMemcpyAsync(result, slot, sizeof(double), hipMemcpyDeviceToHost, S(stream)));
    HPCGX_HIP(hipStreamSynchronize(S(stream)));
  });
}

int hpcgx_waxpby(int64_t n, double alpha, const double *x, double beta, const double *y,
                 double *w, void *stream) {
  return guard([&] {
    if (n < 0) throw Error(HPCGX_EINVAL, "negative length");
    hpcgx::launch_waxpby(n, alpha, x, beta, y, w, S(stream));
  });
}

int hpcgx_tune(hpcgx_problem *p, const char *key, double value) {
  return guard([&] {
    if (!p || !key) throw Error(HPCGX_EINVAL, "NULL argument");
    const std::string k(key);
    hpcgx::Problem &P = p->P;
    P.graph.reset();  // kernel choices may change: recapture on next use
    P.fleet.graph.reset();
    P.mgraph.reset();
    P.bgraph.reset();
    if (k == "col16") {
      P.use_c16 = value != 0.0;
      return;
    }
    if (k == "col8") {
      P.use_c8 = value != 0.0;
      return;
    }
    if (k == "mixed_graph") {
      P.use_mixed_graph = value != 0.0;
      return;
    }
    if (k == "stream_p") {
      P.stream_p = value != 0.0;
      return;
    }
    if (k == "spmv_dot_chunks") {
      P.spmv_dot_chunked = value != 0.0;
      return;
    }
    if (k == "affine8") {
      P.use_affine = value != 0.0;
      return;
    }
    if (k == "rz_fused") {
      P.rz_fused = value != 0.0;
      return;
    }
    if (k == "zero_guess") {
      P.zero_guess = value != 0.0;
      return;
    }
    if (k == "defer_x") {
      P.defer_x = value != 0.0;
      return;
    }
    if (k == "gs_block") {
      if (value != 64 && value != 128 && value != 256 && value != 512)
        throw Error(HPCGX_EINVAL, "gs_block must be 64, 128, 256 or 512");
      P.gs_block = (int)value;
      return;
    }
    if (k == "block_nt") {
      P.block_nt = value != 0.0;
      return;
    }
    if (k == "resid_fused") {
      P.resid_fused = value != 0.0;
      return;
    }
    if (k == "hc_pairs") {
      P.hc_pairs = value != 0.0;
      return;
    }
    if (k == "zero_live") {
      if (value < 0 || value > 15 || value != (int)value)
        throw Error(HPCGX_EINVAL, "zero_live must be an integer from 0 to 15");
      P.zero_live = (int)value;
      return;
    }
    if (k == "block_pairs") {
      P.block_pairs = value != 0.0;
      return;
    }
    if (k == "xr_rows") {
      P.xr_rows = value != 0.0;
      return;
    }
    if (k == "x_pairs") {
      if (value < 0 || value > 15 || value != (int)value)
        throw Error(HPCGX_EINVAL, "x_pairs must be an integer from 0 to 15");
      P.x_pairs = (int)value;
      return;
    }
    if (k == "gs_waves") {
      P.gs_waves = value != 0.0 ? 1 : 0;
      return;
    }
    if (k == "fp16_pairs") {
      P.use_pairs = value != 0.0;
      return;
    }
    if (k == "fleet_blocks") {
      if (value < 0 || value > 65536 || value != (int)value)
        throw Error(HPCGX_EINVAL, "fleet_blocks must be an integer in [0, 65536]");
      P.fleet.grid = (int)value;
      return;
    }
    if (k == "spmv_dot_blocks") {
      if (value < 8 || value > hpcgx::kMaxPartials)
        throw Error(HPCGX_EINVAL, "spmv_dot_blocks must be in [8, 8192]");
      P.spmv_dot_blocks = (int)value / 8 * 8;
      return;
    }
    if (k == "wave") {
      P.use_wave = value != 0.0;
      if (P.use_wave)
        for (int l = 0; l < P.nlevels; ++l)
          if (P.L[l].n >= hpcgx::kWaveMinRows && !P.L[l].wave)
            hpcgx::build_wave(P, P.L[l], hpcgx::kWaveSlabFactor, hpcgx::kWaveRows, nullptr);
      HPCGX_HIP(hipDeviceSynchronize());
      return;
    }
    for (int l = 0; l < P.nlevels; ++l) {
      hpcgx::Level &L = P.L[l];
      if (k == "xcd_remap") {
        L.xcd = value != 0.0;
      } else if (k == "spmv_order") {
        if (L.chunks_interleaved != (value == 0.0)) hpcgx::build_chunks(P, L, value == 0.0);
      } else if (k == "pair_rows") {
        L.pair_rows = value != 0.0;
      } else if (k == "stream_min_bytes") {
        L.stream = (double)L.n * kSlots * 12.0 > value;
      } else if (k == "wave_min_rows") {
        if ((double)L.n >= value) {
          if (!L.wave) hpcgx::build_wave(P, L, hpcgx::kWaveSlabFactor, hpcgx::kWaveRows, nullptr);
        } else {
          hpcgx::drop_wave(P, L);
        }
      } else if (k == "wave_slab_factor") {
        if (value < 1) throw Error(HPCGX_EINVAL, "wave_slab_factor must be >= 1");
        if (L.wave)
          hpcgx::build_wave(P, L, (int)value, L.wave_chunk / hpcgx::kBlock, nullptr);
      } else if (k == "wave_rows") {
        if (value < 1 || value > 16) throw Error(HPCGX_EINVAL, "wave_rows must be in [1, 16]");
        if (L.wave)
          hpcgx::build_wave(P, L, (int)(L.wave_slab / std::max<int64_t>(1, L.bandwidth)),
                            (int)value, nullptr);
      } else {
        throw Error(HPCGX_EINVAL, "unknown tuning key " + k);
      }
    }
    HPCGX_HIP(hipDeviceSynchronize());
  });
}

int hpcgx_autotune(hpcgx_problem *p, int ordering, int iterations, char *report,
                   int report_len) {
  return guard([&] {
    if (!p) throw Error(HPCGX_EINVAL, "NULL problem");
    check_order(ordering);
    if (iterations < 1) throw Error(HPCGX_EINVAL, "iterations must be >= 1");
    hpcgx::Problem &P = p->P;
    const int64_t n = P.L[0].n;
    hpcgx::TempBufD bx(sizeof(double) * 2 * n);
    double *b = static_cast<double *>(bx.p), *x = b + n;
    hpcgx::launch_fill(b, n, 1.0, nullptr);
    // one timed solve of `iterations` PCG iterations (after one warm-up)
    auto timed = [&]() {
      int it;
      double nr, nr0;
      // a fleet is timed with its own CG (per-problem scalars)
      const int Q = std::max(1, P.fleet.count);
      std::vector<int> fit(Q);
      std::vector<double> fnr(Q), fnr0(Q);
      auto solve = [&](int iters) {
        if (P.fleet.count)
          hpcgx::cg_fleet(P, ordering, b, x, iters, 0.0, true, HPCGX_CG_GRAPH, fit.data(),
                          fnr.data(), fnr0.data(), nullptr, nullptr);
        else
          hpcgx::cg(P, ordering, b, x, iters, 0.0, true, 0, &it, &nr, &nr0, nullptr, nullptr);
      };
      hpcgx::launch_fill(x, n, 0.0, nullptr);
      solve(1);
      hpcgx::launch_fill(x, n, 0.0, nullptr);
      HPCGX_HIP(hipDeviceSynchronize());
      const auto t0 = std::chrono::steady_clock::now();
      solve(iterations);
      HPCGX_HIP(hipDeviceSynchronize());
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() /
             iterations;
    };
    struct Knob {
      const char *key;
      std::vector<double> values;
    };
    bool any_c16 = false;
    for (int l = 0; l < P.nlevels; ++l) any_c16 = any_c16 || P.L[l].c16;
    std::vector<Knob> knobs;
    if (any_c16) knobs.push_back({"col8", {1, 0}});
    if (P.fleet.count) {
      knobs.push_back({"fleet_blocks", {0, 512, 1024, 2048, 4096}});
    } else {
      knobs.push_back({"spmv_dot_chunks", {0, 1}});
      knobs.push_back({"spmv_dot_blocks", {1024, 2048, 4096, 8192}});
    }
    knobs.push_back({"spmv_order", {0, 1}});
    knobs.push_back({"xcd_remap", {0, 1}});
    if (P.L[0].c16) {  // the finest colour steps' shape: workgroup size, register cap
      knobs.push_back({"gs_block", {256, 128, 512, 64}});
      knobs.push_back({"gs_waves", {1, 0}});
      knobs.push_back({"x_pairs", {3, 0}});  // applies at 256 threads with the caps
    }
    std::string log;
    char line[160];
    // One knob at a time, the others at their best.  Each candidate is timed
    // in kRounds interleaved rounds (a, b, .., a, b, ..) and keeps its fastest:
    // gains of 1-2% are within the spread of a single short trial.
    constexpr int kRounds = 2;
    for (const Knob &k : knobs) {
      std::vector<double> tmin(k.values.size(), 1e300);
      for (int rd = 0; rd < kRounds; ++rd)
        for (size_t q = 0; q < k.values.size(); ++q) {
          HPCGX_REQUIRE(hpcgx_tune(p, k.key, k.values[q]) == 0, "autotune: tuning call failed");
          tmin[q] = std::min(tmin[q], timed());
        }
      double best_t = 1e300, best_v = k.values[0];
      for (size_t q = 0; q < k.values.size(); ++q) {
        std::snprintf(line, sizeof line, "%s=%g: %.4f ms/iteration\n", k.key, k.values[q],
                      tmin[q] * 1e3);
        log += line;
        if (tmin[q] < best_t) {
          best_t = tmin[q];
          best_v = k.values[q];
        }
      }
      HPCGX_REQUIRE(hpcgx_tune(p, k.key, best_v) == 0, "autotune: tuning call failed");
      std::snprintf(line, sizeof line, "chosen %s=%g\n", k.key, best_v);
      log += line;
    }
    if (report && report_len > 0) {
      std::snprintf(report, (size_t)report_len, "%s", log.c_str());
    }
  });
}

int hpcgx_check_status(hpcgx_problem *p, void *stream) {
  return guard([&] {
    if (!p) throw Error(HPCGX_EINVAL, "NULL problem");
    HPCGX_HIP(hipStreamSynchronize(S(stream)));
    hpcgx::check_status(p->P);
  });
}

int hpcgx_kernel_timing(hpcgx_problem *p, int enable) {
  return guard([&] {
    if (!p) throw Error(HPCGX_EINVAL, "NULL problem");
    if (enable) p->P.klog.used = 0;
    p->P.klog.on = enable != 0;
  });
}

int hpcgx_kernel_timing_read(hpcgx_problem *p, double *total_ms, int64_t *launches) {
  return guard([&] {
    if (!p || !total_ms || !launches) throw Error(HPCGX_EINVAL, "NULL argument");
    auto &K = p->P.klog;
    double t = 0.0;
    if (K.used) HPCGX_HIP(hipEventSynchronize(K.ev[K.used - 1]));
    for (size_t i = 0; i + 1 < K.used; i += 2) {
      float ms = 0.f;
      HPCGX_HIP(hipEventElapsedTime(&ms, K.ev[i], K.ev[i + 1]));
      t += ms;
    }
    *total_ms = t;
    *launches = (int64_t)(K.used / 2);
  });
}

int hpcgx_prepare_mixed(hpcgx_problem *p) {
  return guard([&] {
    if (!p) throw Error(HPCGX_EINVAL, "NULL problem");
    hpcgx::prepare_mixed(p->P);
  });
}

int hpcgx_spmv32(hpcgx_problem *p, int level, const float *x, float *y, void *stream) {
  return guard([&] {
    hpcgx::Level &L = level_of(p, level);
    hpcgx::prepare_mixed(p->P);
    hpcgx::spmv32(p->P, L, x, y, S(stream));
  });
}

int hpcgx_symgs32(hpcgx_problem *p, int level, int ordering, const float *r, float *x,
                  void *stream) {
  return guard([&] {
    level_of(p, level);
    check_order(ordering);
    hpcgx::prepare_mixed(p->P);
    hpcgx::symgs32(p->P, level, ordering, r, x, S(stream));
  });
}

int hpcgx_mg32(hpcgx_problem *p, int level, int ordering, const float *r, float *x,
               void *stream) {
  return guard([&] {
    level_of(p, level);
    check_order(ordering);
    hpcgx::prepare_mixed(p->P);
    hpcgx::mg32(p->P, level, ordering, r, x, S(stream));
  });
}

int hpcgx_cg_mixed(hpcgx_problem *p, int ordering, const double *b, double *x,
                   int max_iter, double tolerance, int recompute, int *niters,
                   double *normr, double *normr0, double *times, void *stream) {
  return guard([&] {
    if (!p || !b || !x || !niters || !normr || !normr0) throw Error(HPCGX_EINVAL, "NULL argument");
    check_order(ordering);
    if (max_iter < 0 || recompute < 0) throw Error(HPCGX_EINVAL, "negative count");
    hpcgx::cg_mixed(p->P, ordering, b, x, max_iter, tolerance, recompute, niters, normr,
                    normr0, times, S(stream));
  });
}

int hpcgx_cg(hpcgx_problem *p, int ordering, const double *b, double *x, int max_iter,
             double tolerance, int precondition, int flags, int *niters, double *normr,
             double *normr0, double *times, void *stream) {
  return guard([&] {
    if (!p || !b || !x || !niters || !normr || !normr0) throw Error(HPCGX_EINVAL, "NULL argument");
    check_order(ordering);
    if (max_iter < 0) throw Error(HPCGX_EINVAL, "negative max_iter");
    hpcgx::cg(p->P, ordering, b, x, max_iter, tolerance, precondition != 0, flags, niters,
              normr, normr0, times, S(stream));
  });
}

int hpcgx_generate_fleet(int count, const int *nx, const int *ny, const int *nz, int nlevels,
                         int flags, hpcgx_problem **out) {
  return guard([&] {
    if (!out || !nx || !ny || !nz) throw Error(HPCGX_EINVAL, "NULL argument");
    *out = nullptr;
    if (count < 1 || count > hpcgx::kMaxFleet)
      throw Error(HPCGX_EINVAL, "fleet size must be in [1, 1024]");
    if (nlevels < 1 || nlevels > hpcgx::kMaxLevels) throw Error(HPCGX_EINVAL, "nlevels out of range");
    const int div = 1 << (nlevels - 1);
    std::vector<int> dims(3 * (size_t)count);
    int64_t total = 0;
    for (int q = 0; q < count; ++q) {
      if (nx[q] < 1 || ny[q] < 1 || nz[q] < 1)
        throw Error(HPCGX_EINVAL, "grid dimensions must be >= 1");
      if (nx[q] % div || ny[q] % div || nz[q] % div)  // GenerateCoarseProblem.cpp:50
        throw Error(HPCGX_EINVAL, "grid dimensions must be divisible by 2^(nlevels-1)");
      dims[3 * q] = nx[q];
      dims[3 * q + 1] = ny[q];
      dims[3 * q + 2] = nz[q];
      total += (int64_t)nx[q] * ny[q] * nz[q];
    }
    if (total >= ((int64_t)1 << 31))
      throw Error(HPCGX_EINVAL, "fleet rows must be < 2^31 (int32 row ids)");
    auto h = std::make_unique<hpcgx_problem>();
    hpcgx::Problem &P = h->P;
    HPCGX_HIP(hipGetDevice(&P.device));
    P.nlevels = nlevels;
    P.generated = true;
    P.L[0].finest = true;
    std::vector<NaturalOwner> nat(nlevels);
    hipStream_t s = nullptr;
    for (int l = 0; l < nlevels; ++l) {
      int64_t n = 0;
      for (int q = 0; q < count; ++q)
        n += (int64_t)(dims[3 * q] >> l) * (dims[3 * q + 1] >> l) * (dims[3 * q + 2] >> l);
      nat[l].alloc(n, (l + 1 < nlevels) ? n / 8 : 0);
      int64_t off = 0, offc = 0;
      for (int q = 0; q < count; ++q) {
        const int x = dims[3 * q] >> l, y = dims[3 * q + 1] >> l, z = dims[3 * q + 2] >> l;
        nat[l].N.parts.push_back(hpcgx::GridPart{off, x, y, z});
        hpcgx::generate_natural_part(nat[l].N, x, y, z, off, offc, s);
        off += (int64_t)x * y * z;
        offc += (int64_t)x * y * z / 8;
      }
      // not one grid: no band order, level dimensions reported as 0
      hpcgx::finalize_level(P, l, nat[l].N, flags, s);
      nat[l].drop_matrix();
    }
    for (int l = 0; l + 1 < nlevels; ++l) hpcgx::link_levels(P, l, nat[l].N, s);
    HPCGX_HIP(hipStreamSynchronize(s));
    nat.clear();
    alloc_solver_state(P);
    hpcgx::prepare_fleet(P, dims, s);
    hpcgx::check_problem(P, nullptr, nullptr, nullptr, s);  // main.cpp:150-159, per problem
    HPCGX_HIP(hipStreamSynchronize(s));
    *out = h.release();
  });
}

int hpcgx_check_problem(const hpcgx_problem *p, const double *b, const double *x,
                        const double *xexact, void *stream) {
  return guard([&] {
    if (!p) throw Error(HPCGX_EINVAL, "NULL problem");
    hpcgx::check_problem(p->P, b, x, xexact, static_cast<hipStream_t>(stream));
  });
}

int hpcgx_fleet_info(const hpcgx_problem *p, int level, int *count, int64_t *row_offsets) {
  return guard([&] {
    if (!p || !count) throw Error(HPCGX_EINVAL, "NULL argument");
    const hpcgx::Problem &P = p->P;
    *count = P.fleet.count;
    if (P.fleet.count == 0) return;
    if (level < 0 || level >= P.nlevels) throw Error(HPCGX_EINVAL, "level out of range");
    if (row_offsets)
      std::copy(P.fleet.off[level].begin(), P.fleet.off[level].end(), row_offsets);
  });
}

int hpcgx_cg_fleet(hpcgx_problem *p, int ordering, const double *b, double *x, int max_iter,
                   double tolerance, int precondition, int flags, int *niters, double *normr,
                   double *normr0, double *times, void *stream) {
  return guard([&] {
    if (!p || !b || !x || !niters || !normr || !normr0) throw Error(HPCGX_EINVAL, "NULL argument");
    check_order(ordering);
    if (max_iter < 0) throw Error(HPCGX_EINVAL, "negative max_iter");
    if (p->P.fleet.count == 0) throw Error(HPCGX_EINVAL, "not a fleet (hpcgx_generate_fleet)");
    hpcgx::cg_fleet(p->P, ordering, b, x, max_iter, tolerance, precondition != 0, flags, niters,
                    normr, normr0, times, S(stream));
  });
}

int hpcgx_spmv_block(hpcgx_problem *p, int level, int s, const double *X, double *Y,
                     void *stream) {
  return guard([&] {
    level_of(p, level);
    if (!X || !Y) throw Error(HPCGX_EINVAL, "NULL vector");
    if (s < 1 || s > 64 || 64 % s) throw Error(HPCGX_EINVAL, "s must divide 64");
    hpcgx::spmv_block(p->P, level, s, X, Y, S(stream));
  });
}

int hpcgx_mg32_block(hpcgx_problem *p, int level, int ordering, int s, const float *R,
                     float *X, void *stream) {
  return guard([&] {
    level_of(p, level);
    check_order(ordering);
    if (!R || !X) throw Error(HPCGX_EINVAL, "NULL vector");
    hpcgx::prepare_block(p->P, s);
    hpcgx::mg32_block(p->P, level, ordering, s, R, X, S(stream));
  });
}

int hpcgx_cg_block_mixed(hpcgx_problem *p, int ordering, int s, const double *B, double *X,
                         int max_iter, double tolerance, int recompute, int *niters,
                         double *normr, double *normr0, double *times, void *stream) {
  return guard([&] {
    if (!p || !B || !X || !niters || !normr || !normr0) throw Error(HPCGX_EINVAL, "NULL argument");
    check_order(ordering);
    if (max_iter < 0) throw Error(HPCGX_EINVAL, "negative max_iter");
    hpcgx::cg_block_mixed(p->P, ordering, s, B, X, max_iter, tolerance, recompute, niters, normr,
                          normr0, times, S(stream));
  });
}

}  // extern "C"
