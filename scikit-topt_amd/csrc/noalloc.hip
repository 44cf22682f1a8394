// Allocation-free percentile: the same radix-select kernels as
// skt_percentile_abs, with their scratch in a static device array instead of
// a buffer the library allocates, for callers that cannot allocate device
// memory (code under graph capture, tight serving loops).
#include <mutex>

#include "common.hpp"

namespace skt {

// Static scratch: selection partials (2 * kMaxBlocks doubles), then the
// 256-bin histogram (room of 128 doubles), then the 16-word selection state;
// 17.5 KB per device, held until the process ends.  Every call on one device
// shares it and it keeps the last call's contents: calls that overlap on one
// device (two streams, or two host threads) race.
constexpr int64_t kNoAllocMaxN = (int64_t)1 << 53;
constexpr int kNoAllocPart = 2 * kMaxBlocks;
constexpr int kNoAllocHist = 256 / 2;
constexpr int kNoAllocDoubles = kNoAllocPart + kNoAllocHist + 16;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static double* g_noalloc_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch`, looked up once per
// device with that device current.  The host-side name `scratch` is only a
// stand-in and never reaches a kernel or a copy.
int noalloc_scratch(double** out) {
  int dev = 0;
  SKT_HIP(hipGetDevice(&dev));
  SKT_CHECK(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    double* ptr = nullptr;
    SKT_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    SKT_CHECK(ptr != nullptr, "hipGetSymbolAddress returned a null scratch address");
    g_noalloc_ptr[dev] = ptr;
  }
  *out = g_noalloc_ptr[dev];
  return SKT_OK;
}

}  // namespace skt

using namespace skt;

extern "C" {

int skt_percentile_abs_noalloc(int64_t n, const double* x, double q, double* out,
                               hipStream_t stream) {
  SKT_CHECK(n >= 1 && x && out, "percentile of an empty array");
  SKT_CHECK(n <= kNoAllocMaxN,
            "skt_percentile_abs_noalloc: n exceeds the maximum of 2^53 entries (the "
            "selection counts are exact float64 integers only up to there)");
  SKT_CHECK(q >= 0.0 && q <= 100.0, "Percentiles must be in the range [0, 100]");
  const unsigned nb = grid_for(n, kBlock, kMaxBlocks);
  SKT_CHECK(2 * (int64_t)nb <= kNoAllocPart,
            "skt_percentile_abs_noalloc: partials exceed the static scratch");
  double* base = nullptr;
  int st = noalloc_scratch(&base);
  if (st) return st;
  return launch_percentile(n, x, q, out, base, (uint32_t*)(base + kNoAllocPart),
                           (uint64_t*)(base + kNoAllocPart + kNoAllocHist), stream);
}

}  // extern "C"
