// Shared helpers for the gfx950 kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/sktopt_amd.h"

namespace skt {

// Last error message of the calling host thread.
void set_error(const std::string& msg);
const char* get_error();

#define SKT_HIP(call)                                                          \
  do {                                                                         \
    hipError_t err_ = (call);                                                  \
    if (err_ != hipSuccess) {                                                  \
      ::skt::set_error(std::string(#call) + ": " + hipGetErrorString(err_));   \
      return SKT_EHIP;                                                         \
    }                                                                          \
  } while (0)

#define SKT_CHECK(cond, msg)                                                   \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::skt::set_error(msg);                                                   \
      return SKT_EINVAL;                                                       \
    }                                                                          \
  } while (0)

#define SKT_LAUNCH_CHECK()                                                     \
  do {                                                                         \
    hipError_t err_ = hipGetLastError();                                       \
    if (err_ != hipSuccess) {                                                  \
      ::skt::set_error(std::string("kernel launch: ") + hipGetErrorString(err_)); \
      return SKT_EHIP;                                                         \
    }                                                                          \
  } while (0)

constexpr int kWave = 64;
constexpr int kBlock = 256;
// Largest grid of the reductions, statistics and percentile (partial slots).
constexpr int kMaxBlocks = 1024;

// np.percentile(|x|, q) ('linear') by radix select, on caller-chosen scratch:
// part >= 2 * min(ceil(n / kBlock), kMaxBlocks) doubles, hist 256 words,
// sel 16 words (design.hip).
int launch_percentile(int64_t n, const double* x, double q, double* out, double* part,
                      uint32_t* hist, uint64_t* sel, hipStream_t stream);

inline unsigned grid_for(int64_t n, int block = kBlock, int64_t cap = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// Wave-level sum (fixed butterfly order, deterministic).
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Wave-level maximum; a NaN on any lane propagates (as numpy's max).
template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, kWave);
    v = (v > o || v != v) ? v : o;
  }
  return v;
}

// Deterministic block sum of one value per thread (blockDim.x == kBlock).
template <typename T>
__device__ inline T block_sum(T v, T* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T s = T(0);
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) s += sh[i];
  }
  return s;  // valid in thread 0 only
}

// Deterministic block maximum (NaN propagates).
template <typename T>
__device__ inline T block_max(T v, T* sh) {
  v = wave_max(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T s = sh[0];
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x / kWave); ++i)
      s = (s > sh[i] || s != s) ? s : sh[i];
  }
  return s;  // valid in thread 0 only
}

// Final fixed-order reduction of per-block partials into out[0] (one block).
template <typename T>
__global__ void reduce_partials_sum(const T* __restrict__ part, int n, T* __restrict__ out) {
  __shared__ T sh[kBlock / kWave];
  T v = T(0);
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
  T s = block_sum(v, sh);
  if (threadIdx.x == 0) out[0] = s;
}

// numpy.clip semantics (NaN propagates): min(max(x, lo), hi).
__device__ __forceinline__ double clip_np(double x, double lo, double hi) {
  double y = x < lo ? lo : x;
  return y > hi ? hi : y;
}

// OC multiplicative update of one design element at multiplier lmid
// (core/optimizers/oc.py:51-65): rate = clip((-dC / (lmid + eps))^eta,
// rate_min, rate_max); candidate = clip(rho * rate, max(rho - move, rho_min),
// min(rho + move, rho_max)), in the reference's operation order.
__device__ __forceinline__ double oc_candidate(double dC, double re, double lmid, double eps,
                                               double eta, double rate_min, double rate_max,
                                               double move_limit, double rho_min,
                                               double rho_max) {
  double s = -dC;
  s = s / (lmid + eps);
  s = pow(s, eta);
  s = clip_np(s, rate_min, rate_max);
  double lo = re - move_limit;
  lo = (lo < rho_min) ? rho_min : lo;     // np.maximum
  double hi = re + move_limit;
  hi = (hi > rho_max) ? rho_max : hi;     // np.minimum
  return clip_np(re * s, lo, hi);
}

}  // namespace skt
