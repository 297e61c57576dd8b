// noalloc.hip -- hpcgx_dot_noalloc: the dot product of hpcgx_dot for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc), with no problem handle.
//
// hpcgx_dot keeps its reduction scratch (the partials of its two passes and
// the completion ticket of the last-workgroup finish) in the problem's arena.
// Here the same scratch is a file-scope device array, one copy per GPU for
// the life of the process.  Its address is looked up once per device with
// hipGetSymbolAddress (the array's name is only a host-side stand-in in host
// code) and kept; every launch takes that address.  The kernels are
// hpcgx_dot's own (launch_dot_raw), so the result is hpcgx_dot's bit for bit.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (on different streams, or from several threads) race on it.
#include <mutex>

#include "internal.hpp"

namespace {
// hpcgx_dot's scratch: the first pass's partials (kDotParts), the partials
// of the last-workgroup finish (at most kReduceBlocks workgroups) and the
// ticket.  Those bounds hold for every n (longer vectors take the
// grid-stride reduction, as in hpcgx_dot), so no length is refused.
constexpr size_t kPartialsOffset = (size_t)hpcgx::kDotParts * sizeof(double);
constexpr size_t kTicketOffset = kPartialsOffset + (size_t)hpcgx::kReduceBlocks * sizeof(double);
constexpr size_t kScratchBytes = kTicketOffset + 256;  // 1 MiB + 32 KiB + the ticket
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_mu;
void *g_scratch[kMaxDevices];  // the array's device address, per device

// The current device's copy of `scratch`, looked up on the first call on
// that device (with it current) and kept.
unsigned char *device_scratch() {
  int dev = -1;
  HPCGX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    throw hpcgx::Error(HPCGX_EUNSUP, "device ordinal beyond the no-allocation scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    HPCGX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw hpcgx::Error(HPCGX_EHIP, "no device address for the static scratch");
    g_scratch[dev] = p;
  }
  return static_cast<unsigned char *>(g_scratch[dev]);
}
}  // namespace

extern "C" int hpcgx_dot_noalloc(int64_t n, const double *x, const double *y,
                                 double *result_dev, void *stream) {
  return hpcgx::guard([&] {
    if (!result_dev) throw hpcgx::Error(HPCGX_EINVAL, "NULL argument");
    if (n < 0) throw hpcgx::Error(HPCGX_EINVAL, "negative length");
    if (n > 0 && (!x || !y)) throw hpcgx::Error(HPCGX_EINVAL, "NULL vector");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned char *base = device_scratch();
    double *parts = reinterpret_cast<double *>(base);
    double *partials = reinterpret_cast<double *>(base + kPartialsOffset);
    unsigned *ticket = reinterpret_cast<unsigned *>(base + kTicketOffset);
    // The last-workgroup finish needs the ticket at 0.  Each launch leaves it
    // at 0, but a launch that never ran to its end would not: clear it with a
    // kernel (no memset node, so the call can be captured in a graph).
    hpcgx::launch_zero_u32(ticket, 1, s);
    hpcgx::launch_dot_raw(n, x, y, parts, partials, ticket, nullptr, result_dev,
                          hpcgx::kFinishStore, s);
  });
}
