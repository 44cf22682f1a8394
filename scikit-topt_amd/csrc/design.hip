// Design-variable kernels (float64): projection, interpolation, compliance
// sensitivity, LogMOC update, reductions and the exact percentile.
#include <cmath>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"

namespace skt {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* get_error() { return g_last_error.c_str(); }

// numpy.clip semantics (NaN propagates): min(max(x, lo), hi).
__device__ __forceinline__ double clipd(double x, double lo, double hi) {
  double y = x < lo ? lo : x;
  return y > hi ? hi : y;
}

// ---------------------------------------------------------------- transforms

__global__ void k_heaviside(int64_t n, const double* __restrict__ x, double beta,
                            double eta, double tbe, double denom,
                            double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double t = (x[i] - eta) * beta;
    t = tanh(t);
    out[i] = (tbe + t) / denom;
  }
}

__global__ void k_heaviside_deriv(int64_t n, const double* __restrict__ x,
                                  double beta, double eta, double denom,
                                  double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double t = (x[i] - eta) * beta;
    double c = cosh(t);
    double s = 1.0 / (c * c);
    s = s * beta;
    out[i] = s / denom;
  }
}

template <typename TO>
__global__ void k_interp(int kind, int64_t n, const double* __restrict__ rho,
                         double E0, double Emin, double p, TO* __restrict__ out) {
  const double dE = E0 - Emin;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double r = rho[i];
    double v;
    if (kind == SKT_SIMP) {
      v = Emin + dE * pow(r, p);
    } else {
      v = Emin + dE * (r / (1.0 + p * (1.0 - r)));
    }
    out[i] = (TO)v;
  }
}

__device__ __forceinline__ double dC_value(int kind, double rho, double U,
                                           double E0, double Emin, double p) {
  const double dEm = E0 - Emin;
  if (kind == SKT_SIMP) {
    double r = rho > 1e-6 ? rho : 1e-6;
    if (rho != rho) r = rho;
    double dE = (p * dEm) * pow(r, p - 1.0);
    double E = Emin + dEm * pow(r, p);
    double Ec = E > 1e-12 ? E : 1e-12;
    if (E != E) Ec = E;
    return ((-2.0 * U) * dE) / Ec;
  } else {
    // SKT_RAMP: the reference's expression (core/derivatives.py:56-61);
    // SKT_RAMP_EXACT: the derivative of rho / (1 + p (1 - rho)), (1 + p) / den^2
    double den = 1.0 + p * (1.0 - rho);
    double dE = kind == SKT_RAMP ? dEm * (den - p * rho) / (den * den)
                                 : dEm * (1.0 + p) / (den * den);
    double E = Emin + dEm * (rho / den);
    double Ec = E > 1e-12 ? E : 1e-12;
    if (E != E) Ec = E;
    return ((-2.0 * U) * dE) / Ec;
  }
}

__global__ void k_dC(int kind, int64_t n, const double* __restrict__ rho,
                     const double* __restrict__ U, int64_t stride, double E0,
                     double Emin, double p, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    out[i] = dC_value(kind, rho[i], U[i * stride], E0, Emin, p);
  }
}

__global__ void k_sens_chain(int kind, int64_t n, int ncols,
                             const double* __restrict__ rho_proj,
                             const double* __restrict__ energy,
                             const double* __restrict__ extra,
                             const double* __restrict__ rho_filt, double beta,
                             double eta, double denom, double E0, double Emin,
                             double p, double scale, double* __restrict__ g) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double t = (rho_filt[i] - eta) * beta;
    double c = cosh(t);
    double dH = ((1.0 / (c * c)) * beta) / denom;
    double r = rho_proj[i];
    double acc = 0.0;
    for (int k = 0; k < ncols; ++k) {
      double d = dC_value(kind, r, energy[i * ncols + k], E0, Emin, p);
      if (extra) d = d + extra[i * ncols + k];
      acc += d * dH;
    }
    g[i] = acc * scale;
  }
}

// One chain product per load case, g (n, ncols): for filters whose adjoint
// is applied per load case (the nodal Helmholtz gradient clamps each call).
__global__ void k_sens_chain_cols(int kind, int64_t n, int ncols,
                                  const double* __restrict__ rho_proj,
                                  const double* __restrict__ energy,
                                  const double* __restrict__ extra,
                                  const double* __restrict__ rho_filt, double beta,
                                  double eta, double denom, double E0, double Emin,
                                  double p, double* __restrict__ g) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double t = (rho_filt[i] - eta) * beta;
    double c = cosh(t);
    double dH = ((1.0 / (c * c)) * beta) / denom;
    double r = rho_proj[i];
    for (int k = 0; k < ncols; ++k) {
      double d = dC_value(kind, r, energy[i * ncols + k], E0, Emin, p);
      if (extra) d = d + extra[i * ncols + k];
      g[i * ncols + k] = d * dH;
    }
  }
}

__global__ void k_design_gather(int64_t nd, const int32_t* __restrict__ ids,
                                const double* __restrict__ full, double* __restrict__ d) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nd;
       r += (int64_t)gridDim.x * blockDim.x)
    d[r] = full[ids[r]];
}

__global__ void k_design_scatter(int64_t nd, const int32_t* __restrict__ ids,
                                 const double* __restrict__ d, double* __restrict__ full) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nd;
       r += (int64_t)gridDim.x * blockDim.x)
    full[ids[r]] = d[r];
}

__global__ void k_oc_candidate_at(int64_t nd, const int32_t* __restrict__ ids,
                                  const double* __restrict__ rho_e,
                                  const double* __restrict__ dC, double lmid, double eps,
                                  double eta, double rate_min, double rate_max,
                                  double move_limit, double rho_min, double rho_max,
                                  double* __restrict__ cand_full, double* __restrict__ rho_out) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nd;
       r += (int64_t)gridDim.x * blockDim.x) {
    const double c = oc_candidate(dC[r], rho_e[r], lmid, eps, eta, rate_min, rate_max,
                                  move_limit, rho_min, rho_max);
    rho_out[r] = c;
    if (cand_full) cand_full[ids[r]] = c;
  }
}

__global__ void k_logmoc(int64_t n, double* __restrict__ rho,
                         const double* __restrict__ dL, double eta,
                         double move_limit, double rmin, double rmax,
                         double lclip) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double step = clipd(dL[i], -lclip, lclip);
    double r = clipd(rho[i], rmin, rmax);
    double lo = log(r);
    double up = exp(lo);
    up = move_limit / up;
    up = up + 1.0;
    up = log(up);
    lo = lo - up;
    up = lo + 2.0 * up;
    double x = log(r);
    x = x - eta * step;
    x = clipd(x, lo, up);
    x = exp(x);
    rho[i] = clipd(x, rmin, rmax);
  }
}

// out = a*x + b*y (op 0, y may be null) or out = a*x*y (op 1)
__global__ void k_ewise(int op, int64_t n, double a, const double* __restrict__ x, double b,
                        const double* __restrict__ y, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (op == 0) out[i] = y ? a * x[i] + b * y[i] : a * x[i];
    else out[i] = (a * x[i]) * y[i];
  }
}

// ---------------------------------------------------------------- reductions

// Scratch for reduction partials and selection state, one set per
// (device, stream).  Calls on one stream run in stream order, so they may
// share a set; calls on different streams get different sets and cannot
// overwrite each other's partials.  A set is allocated the first time its
// stream is seen and kept until the process ends (a destroyed stream's
// handle may be reused by a new stream, which then inherits the set).
struct Scratch {
  double* part = nullptr;      // kMaxBlocks * 64 doubles
  uint32_t* hist = nullptr;    // 256 bins
  uint64_t* sel = nullptr;     // selection state
};
static std::mutex g_scratch_mu;
static std::map<std::pair<int, uintptr_t>, Scratch> g_scratch;

int get_scratch(hipStream_t stream, Scratch** out) {
  int dev = 0;
  SKT_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  Scratch& s = g_scratch[std::make_pair(dev, (uintptr_t)stream)];
  if (!s.part) {
    SKT_HIP(hipMalloc(&s.part, sizeof(double) * kMaxBlocks * 64));
    SKT_HIP(hipMalloc(&s.hist, sizeof(uint32_t) * 256));
    SKT_HIP(hipMalloc(&s.sel, sizeof(uint64_t) * 16));
  }
  *out = &s;
  return SKT_OK;
}

__global__ void k_reduce(int op, int64_t n, const double* __restrict__ x,
                         const double* __restrict__ y, double* __restrict__ part) {
  __shared__ double sh[kBlock / kWave];
  double v = (op == SKT_RED_SUM || op == SKT_RED_DOT) ? 0.0 : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    switch (op) {
      case SKT_RED_SUM: v += x[i]; break;
      case SKT_RED_DOT: v += x[i] * y[i]; break;
      case SKT_RED_MAXABS: { double a = fabs(x[i]); v = (a > v || a != a) ? a : v; } break;
      default: { double a = fabs(x[i] - y[i]); v = (a > v || a != a) ? a : v; } break;
    }
  }
  double r = (op == SKT_RED_SUM || op == SKT_RED_DOT) ? block_sum(v, sh) : block_max(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void k_reduce_final(int op, int nb, const double* __restrict__ part,
                               double* __restrict__ out) {
  __shared__ double sh[kBlock / kWave];
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    double a = part[i];
    if (op == SKT_RED_SUM || op == SKT_RED_DOT) v += a;
    else v = (a > v || a != a) ? a : v;
  }
  double r = (op == SKT_RED_SUM || op == SKT_RED_DOT) ? block_sum(v, sh) : block_max(v, sh);
  if (threadIdx.x == 0) out[0] = r;
}

// Column-wise reductions: block b handles a strided slice of rows; each
// thread owns column (tid % ncols_p) where ncols_p is a power of two.
__global__ void k_reduce_cols(int op, int64_t n, int ncols, int cp,
                              const double* __restrict__ x,
                              const double* __restrict__ y,
                              double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int c = threadIdx.x % cp;
  const int rlane = threadIdx.x / cp;
  const int rows_per_block = blockDim.x / cp;
  double v = 0.0;
  if (c < ncols) {
    for (int64_t i = blockIdx.x * (int64_t)rows_per_block + rlane; i < n;
         i += (int64_t)gridDim.x * rows_per_block) {
      double a = x[i * ncols + c];
      if (op == SKT_RED_DOT) v += a * y[i * ncols + c];
      else if (op == SKT_RED_SUM) v += a;
      else { a = fabs(a); v = (a > v || a != a) ? a : v; }
    }
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s >= cp; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double a = sh[threadIdx.x], b = sh[threadIdx.x + s];
      sh[threadIdx.x] = (op != SKT_RED_MAXABS) ? a + b : ((b > a || b != b) ? b : a);
    }
    __syncthreads();
  }
  if ((int)threadIdx.x < ncols) part[(int64_t)blockIdx.x * ncols + threadIdx.x] = sh[threadIdx.x];
}

__global__ void k_reduce_cols_final(int op, int nb, int ncols,
                                    const double* __restrict__ part,
                                    double* __restrict__ out) {
  __shared__ double sh[kBlock / kWave];
  for (int c = 0; c < ncols; ++c) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      double a = part[(int64_t)i * ncols + c];
      if (op != SKT_RED_MAXABS) v += a;
      else v = (a > v || a != a) ? a : v;
    }
    double r = (op != SKT_RED_MAXABS) ? block_sum(v, sh) : block_max(v, sh);
    if (threadIdx.x == 0) out[c] = r;
    __syncthreads();
  }
}

// Row means of a (n, k) row-major block, k <= 128, summed as numpy's
// pairwise sum does for short rows (sequential below 8 values; else eight
// partial sums combined pairwise, then the tail), divided by k.
__global__ void k_row_mean(int64_t n, int k, const double* __restrict__ x,
                           double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double* a = x + i * k;
    double res;
    if (k < 8) {
      res = 0.0;
      for (int j = 0; j < k; ++j) res += a[j];
    } else {
      double r[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) r[j] = a[j];
      const int body = k - (k % 8);
      for (int b = 8; b < body; b += 8)
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] += a[b + j];
      res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
      for (int j = body; j < k; ++j) res += a[j];
    }
    out[i] = res / (double)k;
  }
}

// ------------------------------------------------------------------- stats
// [min, mean, max, population std] of x.  Block b takes the contiguous
// chunk [b*chunk, min(n, (b+1)*chunk)), whose count is c_b.  A first pass
// gives its sum, min, max and the shift a_b = sum / c_b; a second pass over
// the chunk (in cache) gives d1_b = sum (x - a_b) and d2_b = sum (x - a_b)^2.
// The final block pools the chunks exactly about the global mean m:
// M2 = sum_b d2_b + 2 (a_b - m) d1_b + c_b (a_b - m)^2, which holds for any
// shift, so the rounding of a_b does not enter the result.
constexpr int kStatsChunk = 4 * kBlock;
constexpr int kStatsPart = 8;   // c, sum, a, d1, d2, min, max, (pad)

__global__ void k_stats_part(int64_t n, int64_t chunk, const double* __restrict__ x,
                             double* __restrict__ part) {
  __shared__ double sh[kBlock / kWave];
  __shared__ double shift_sh;
  const int64_t a = (int64_t)blockIdx.x * chunk;
  const int64_t b = a + chunk < n ? a + chunk : n;
  double s = 0.0, hi = -INFINITY, nlo = -INFINITY;
  for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) {
    const double v = x[i];
    s += v;
    hi = (v > hi || v != v) ? v : hi;
    nlo = (-v > nlo || v != v) ? -v : nlo;
  }
  const double sum = block_sum(s, sh);
  if (threadIdx.x == 0) shift_sh = sum / (double)(b - a);
  hi = block_max(hi, sh);           // its barriers publish shift_sh
  nlo = block_max(nlo, sh);
  __syncthreads();
  const double m = shift_sh;
  double d1 = 0.0, d2 = 0.0;
  for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) {
    const double d = x[i] - m;
    d1 += d;
    d2 += d * d;
  }
  d1 = block_sum(d1, sh);
  d2 = block_sum(d2, sh);
  if (threadIdx.x == 0) {
    double* p = part + kStatsPart * (int64_t)blockIdx.x;
    p[0] = (double)(b - a);
    p[1] = sum;
    p[2] = m;
    p[3] = d1;
    p[4] = d2;
    p[5] = -nlo;
    p[6] = hi;
  }
}

__global__ void k_stats_final(int nb, int64_t n, const double* __restrict__ part,
                              double* __restrict__ out) {
  __shared__ double sh[kBlock / kWave];
  __shared__ double mean_sh;
  double s = 0.0, hi = -INFINITY, nlo = -INFINITY;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const double* p = part + kStatsPart * i;
    s += p[1];
    hi = (p[6] > hi || p[6] != p[6]) ? p[6] : hi;
    nlo = (-p[5] > nlo || p[5] != p[5]) ? -p[5] : nlo;
  }
  const double sum = block_sum(s, sh);
  if (threadIdx.x == 0) mean_sh = sum / (double)n;
  hi = block_max(hi, sh);
  nlo = block_max(nlo, sh);
  __syncthreads();
  const double m = mean_sh;
  double m2 = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const double* p = part + kStatsPart * i;
    const double e = p[2] - m;
    m2 += p[4] + 2.0 * e * p[3] + p[0] * e * e;
  }
  const double M2 = block_sum(m2, sh);
  if (threadIdx.x == 0) {
    out[0] = -nlo;
    out[1] = m;
    out[2] = hi;
    out[3] = sqrt((M2 < 0.0 ? 0.0 : M2) / (double)n);  // NaN stays NaN
  }
}

// ---------------------------------------------------------------- percentile
// Exact k-th smallest of |x| by an 8-bit MSB-first radix select on the
// float64 bit patterns (non-negative doubles order like their uint64 bits).
// sel state: [0] prefix, [1] remaining rank k, [2] value k-th, [3] count <= v,
//            [4] next larger value bits.

__global__ void k_sel_init(uint64_t* sel, uint32_t* hist, int64_t k) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
  if (threadIdx.x == 0) {
    sel[0] = 0;
    sel[1] = (uint64_t)k;
  }
}

__global__ void k_sel_hist(int64_t n, const double* __restrict__ x, int shift,
                           const uint64_t* __restrict__ sel,
                           uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const uint64_t prefix = sel[0];
  const uint64_t hmask = (shift >= 56) ? 0ull : (~0ull << (shift + 8));
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t b = (uint64_t)__double_as_longlong(fabs(x[i]));
    if ((b & hmask) == prefix) atomicAdd(&h[(b >> shift) & 0xff], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

__global__ void k_sel_pick(int shift, uint64_t* sel, uint32_t* hist) {
  if (threadIdx.x == 0) {
    uint64_t k = sel[1];
    uint64_t cum = 0;
    int bin = 255;
    for (int i = 0; i < 256; ++i) {
      if (cum + hist[i] > k) { bin = i; break; }
      cum += hist[i];
    }
    sel[0] |= ((uint64_t)bin << shift);
    sel[1] = k - cum;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
}

// Count of |x| <= v and the smallest |x| > v (bits), per block.
__global__ void k_sel_next(int64_t n, const double* __restrict__ x,
                           const uint64_t* __restrict__ sel,
                           double* __restrict__ part) {
  __shared__ double sh[kBlock / kWave];
  const double v = __longlong_as_double((long long)sel[0]);
  double cnt = 0.0;
  double nxt = INFINITY;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double a = fabs(x[i]);
    if (a <= v) cnt += 1.0;
    else if (a < nxt) nxt = a;
  }
  double c = block_sum(cnt, sh);
  __syncthreads();
  double m = -block_max(-nxt, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = c;
    part[2 * blockIdx.x + 1] = m;
  }
}

__global__ void k_sel_final(int nb, int64_t k, double gamma, int take_max,
                            const uint64_t* __restrict__ sel,
                            const double* __restrict__ part,
                            double* __restrict__ out) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  double cnt = 0.0, nxt = INFINITY;
  for (int i = 0; i < nb; ++i) {
    cnt += part[2 * i];
    double m = part[2 * i + 1];
    if (m < nxt) nxt = m;
  }
  const double a = __longlong_as_double((long long)sel[0]);
  // a[k+1] equals a[k] when more than k+1 entries are <= a[k].
  double b = (cnt >= (double)(k + 2)) ? a : nxt;
  if (take_max) { out[0] = a; return; }
  // numpy _lerp: a + (b-a)*t, or b - (b-a)*(1-t) where t >= 0.5
  double diff = b - a;
  double r = (gamma >= 0.5) ? (b - diff * (1.0 - gamma)) : (a + diff * gamma);
  out[0] = r;
}

// OC sensitivity normalisation (core/optimizers/oc.py:187-196):
// s = max(scale, floor); rs = first ? s : keep*rs + (1-keep)*s; out = dC / rs
__global__ void k_oc_scale(int64_t n, const double* __restrict__ dC,
                           const double* __restrict__ scale, double floor_,
                           double keep, double one_minus_keep, int first,
                           double* __restrict__ rs, double* __restrict__ out) {
  __shared__ double r;
  if (threadIdx.x == 0) {
    double s = scale[0];
    s = s > floor_ ? s : floor_;
    r = first ? s : keep * rs[0] + one_minus_keep * s;
  }
  __syncthreads();
  const double rv = r;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = dC[i] / rv;
}

__global__ void k_store_scalar(const double* __restrict__ scale, double floor_,
                               double keep, double one_minus_keep, int first,
                               double* __restrict__ rs) {
  double s = scale[0];
  s = s > floor_ ? s : floor_;
  rs[0] = first ? s : keep * rs[0] + one_minus_keep * s;
}

// KKT residual (core/optimizers/oc.py:230-240): max over rho in
// (rmin+1e-6, rmax-1e-6) of |dC_raw + (lmid*rs) dV|; 0 when the set is empty.
__global__ void k_oc_kkt(int64_t n, const double* __restrict__ dC,
                         const double* __restrict__ rho, const double* __restrict__ dV,
                         const double* __restrict__ res, const double* __restrict__ rs,
                         double lo, double hi, double* __restrict__ part) {
  __shared__ double sh[kBlock / kWave];
  const double f = res[0] * rs[0];
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double r = rho[i];
    if (r > lo && r < hi) {
      const double a = fabs(dC[i] + f * dV[i]);
      v = (a > v || a != a) ? a : v;
    }
  }
  double m = block_max(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// Percentile of |x| (numpy 'linear') by the radix select above, on caller-
// chosen scratch: part >= 2 * grid partials, hist 256 bins, sel 16 words.
// Every scratch word is written by a kernel before it is read (k_sel_init
// clears the histogram), so the scratch needs no clearing by the host.
int launch_percentile(int64_t n, const double* x, double q, double* out, double* part,
                      uint32_t* hist, uint64_t* sel, hipStream_t stream) {
  // numpy 'linear': virtual index (n-1)*q/100, floor, gamma, bounds.
  const double qq = q / 100.0;
  const double vidx = (double)(n - 1) * qq;
  int take_max = vidx >= (double)(n - 1);
  double prev = std::floor(vidx);
  double gamma = vidx - prev;
  int64_t k = take_max ? (n - 1) : (int64_t)prev;
  if (vidx < 0) { k = 0; gamma = 0.0; }
  hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(256), 0, stream, sel, hist, k);
  unsigned nb = grid_for(n, kBlock, kMaxBlocks);
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_sel_hist, dim3(nb), dim3(kBlock), 0, stream, n, x, shift, sel, hist);
    hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(kBlock), 0, stream, shift, sel, hist);
  }
  hipLaunchKernelGGL(k_sel_next, dim3(nb), dim3(kBlock), 0, stream, n, x, sel, part);
  hipLaunchKernelGGL(k_sel_final, dim3(1), dim3(64), 0, stream, (int)nb, k, gamma, take_max,
                     sel, part, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

}  // namespace skt

using namespace skt;

extern "C" {

const char* skt_last_error(void) { return skt::get_error(); }
int skt_abi_version(void) { return 1; }

int skt_heaviside_project(int64_t n, const double* x, double beta, double eta,
                          double* out, hipStream_t stream) {
  SKT_CHECK(n >= 0, "n must be >= 0");
  if (n == 0) return SKT_OK;
  SKT_CHECK(x && out, "null pointer");
  const double tbe = std::tanh(beta * eta);
  const double denom = std::tanh(beta * eta) + std::tanh(beta * (1.0 - eta)) + 1e-12;
  hipLaunchKernelGGL(k_heaviside, dim3(grid_for(n)), dim3(kBlock), 0, stream,
                     n, x, beta, eta, tbe, denom, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_heaviside_derivative(int64_t n, const double* x, double beta,
                             double eta, double* out, hipStream_t stream) {
  SKT_CHECK(n >= 0, "n must be >= 0");
  if (n == 0) return SKT_OK;
  SKT_CHECK(x && out, "null pointer");
  const double denom = std::tanh(beta * eta) + std::tanh(beta * (1.0 - eta)) + 1e-12;
  hipLaunchKernelGGL(k_heaviside_deriv, dim3(grid_for(n)), dim3(kBlock), 0,
                     stream, n, x, beta, eta, denom, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_interpolate(int kind, int64_t n, const double* rho, double E0,
                    double Emin, double p, void* out, int out_dtype,
                    hipStream_t stream) {
  SKT_CHECK(kind == SKT_SIMP || kind == SKT_RAMP || kind == SKT_RAMP_EXACT,
            "Interpolation method must be SIMP or RAMP.");
  SKT_CHECK(n >= 0, "n must be >= 0");
  if (n == 0) return SKT_OK;
  SKT_CHECK(rho && out, "null pointer");
  if (out_dtype == SKT_F64) {
    hipLaunchKernelGGL(k_interp<double>, dim3(grid_for(n)), dim3(kBlock), 0,
                       stream, kind, n, rho, E0, Emin, p, (double*)out);
  } else if (out_dtype == SKT_F32) {
    hipLaunchKernelGGL(k_interp<float>, dim3(grid_for(n)), dim3(kBlock), 0,
                       stream, kind, n, rho, E0, Emin, p, (float*)out);
  } else {
    SKT_CHECK(false, "out_dtype must be SKT_F64 or SKT_F32");
  }
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_dC_drho(int kind, int64_t n, const double* rho, const double* energy,
                int64_t energy_stride, double E0, double Emin, double p,
                double* out, hipStream_t stream) {
  SKT_CHECK(kind == SKT_SIMP || kind == SKT_RAMP || kind == SKT_RAMP_EXACT,
            "Interpolation method must be SIMP or RAMP.");
  SKT_CHECK(n >= 0 && energy_stride >= 1, "bad size/stride");
  if (n == 0) return SKT_OK;
  SKT_CHECK(rho && energy && out, "null pointer");
  hipLaunchKernelGGL(k_dC, dim3(grid_for(n)), dim3(kBlock), 0, stream, kind, n,
                     rho, energy, energy_stride, E0, Emin, p, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_sensitivity_chain(int kind, int64_t n, int ncols,
                          const double* rho_proj, const double* energy,
                          const double* rho_filt, double beta, double eta,
                          double E0, double Emin, double p, double scale, double* g,
                          hipStream_t stream) {
  SKT_CHECK(kind == SKT_SIMP || kind == SKT_RAMP || kind == SKT_RAMP_EXACT,
            "Interpolation method must be SIMP or RAMP.");
  SKT_CHECK(n >= 0 && ncols >= 1, "bad size");
  if (n == 0) return SKT_OK;
  SKT_CHECK(rho_proj && energy && rho_filt && g, "null pointer");
  const double denom = std::tanh(beta * eta) + std::tanh(beta * (1.0 - eta)) + 1e-12;
  hipLaunchKernelGGL(k_sens_chain, dim3(grid_for(n)), dim3(kBlock), 0, stream,
                     kind, n, ncols, rho_proj, energy, (const double*)nullptr, rho_filt, beta,
                     eta, denom, E0, Emin, p, scale, g);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_sensitivity_chain_cols(int kind, int64_t n, int ncols, const double* rho_proj,
                               const double* energy, const double* rho_filt, double beta,
                               double eta, double E0, double Emin, double p, double* g,
                               hipStream_t stream) {
  SKT_CHECK(kind == SKT_SIMP || kind == SKT_RAMP || kind == SKT_RAMP_EXACT,
            "Interpolation method must be SIMP or RAMP.");
  SKT_CHECK(n >= 0 && ncols >= 1, "bad size");
  if (n == 0) return SKT_OK;
  SKT_CHECK(rho_proj && energy && rho_filt && g, "null pointer");
  const double denom = std::tanh(beta * eta) + std::tanh(beta * (1.0 - eta)) + 1e-12;
  hipLaunchKernelGGL(k_sens_chain_cols, dim3(grid_for(n)), dim3(kBlock), 0, stream, kind, n,
                     ncols, rho_proj, energy, (const double*)nullptr, rho_filt, beta, eta, denom,
                     E0, Emin, p, g);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_sensitivity_chain_extra(int kind, int64_t n, int ncols, const double* rho_proj,
                                const double* energy, const double* extra,
                                const double* rho_filt, double beta, double eta, double E0,
                                double Emin, double p, double scale, int per_column, double* g,
                                hipStream_t stream) {
  SKT_CHECK(kind == SKT_SIMP || kind == SKT_RAMP || kind == SKT_RAMP_EXACT,
            "Interpolation method must be SIMP or RAMP.");
  SKT_CHECK(n >= 0 && ncols >= 1, "bad size");
  if (n == 0) return SKT_OK;
  SKT_CHECK(rho_proj && energy && rho_filt && g, "null pointer");
  const double denom = std::tanh(beta * eta) + std::tanh(beta * (1.0 - eta)) + 1e-12;
  if (per_column)
    hipLaunchKernelGGL(k_sens_chain_cols, dim3(grid_for(n)), dim3(kBlock), 0, stream, kind, n,
                       ncols, rho_proj, energy, extra, rho_filt, beta, eta, denom, E0, Emin, p,
                       g);
  else
    hipLaunchKernelGGL(k_sens_chain, dim3(grid_for(n)), dim3(kBlock), 0, stream, kind, n, ncols,
                       rho_proj, energy, extra, rho_filt, beta, eta, denom, E0, Emin, p, scale,
                       g);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_design_gather(int64_t n_design, const int32_t* design_ids, const double* full,
                      double* design, hipStream_t stream) {
  SKT_CHECK(n_design >= 0, "n_design must be >= 0");
  if (n_design == 0) return SKT_OK;
  SKT_CHECK(design_ids && full && design, "null pointer");
  hipLaunchKernelGGL(k_design_gather, dim3(grid_for(n_design)), dim3(kBlock), 0, stream,
                     n_design, design_ids, full, design);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_design_scatter(int64_t n_design, const int32_t* design_ids, const double* design,
                       double* full, hipStream_t stream) {
  SKT_CHECK(n_design >= 0, "n_design must be >= 0");
  if (n_design == 0) return SKT_OK;
  SKT_CHECK(design_ids && full && design, "null pointer");
  hipLaunchKernelGGL(k_design_scatter, dim3(grid_for(n_design)), dim3(kBlock), 0, stream,
                     n_design, design_ids, design, full);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_oc_candidate(int64_t n_design, const int32_t* design_ids, const double* rho_e,
                     const double* dC, double lmid, double eta, double rho_min,
                     double rho_max, double rate_min, double rate_max, double move_limit,
                     double* rho_full, double* rho_out, hipStream_t stream) {
  SKT_CHECK(n_design >= 0, "n_design must be >= 0");
  if (n_design == 0) return SKT_OK;
  SKT_CHECK(rho_e && dC && rho_out && (!rho_full || design_ids), "null pointer");
  hipLaunchKernelGGL(k_oc_candidate_at, dim3(grid_for(n_design)), dim3(kBlock), 0, stream,
                     n_design, design_ids, rho_e, dC, lmid, 1e-12, eta, rate_min, rate_max,
                     move_limit, rho_min, rho_max, rho_full, rho_out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_logmoc_update(int64_t n, double* rho, const double* dL, double eta,
                      double move_limit, double rho_min, double rho_max,
                      double clip, hipStream_t stream) {
  SKT_CHECK(n >= 0, "n must be >= 0");
  if (n == 0) return SKT_OK;
  SKT_CHECK(rho && dL, "null pointer");
  hipLaunchKernelGGL(k_logmoc, dim3(grid_for(n)), dim3(kBlock), 0, stream, n,
                     rho, dL, eta, move_limit, rho_min, rho_max, clip);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_reduce(int op, int64_t n, const double* x, const double* y,
               double* out, hipStream_t stream) {
  SKT_CHECK(op >= SKT_RED_SUM && op <= SKT_RED_MAXABSDIFF, "bad reduction op");
  SKT_CHECK(n >= 0 && x && out, "bad argument");
  SKT_CHECK(!(op == SKT_RED_DOT || op == SKT_RED_MAXABSDIFF) || y, "y required");
  Scratch* s;
  int st = get_scratch(stream, &s);
  if (st) return st;
  unsigned nb = grid_for(n, kBlock, kMaxBlocks);
  hipLaunchKernelGGL(k_reduce, dim3(nb), dim3(kBlock), 0, stream, op, n, x, y,
                     s->part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, stream, op,
                     (int)nb, s->part, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_reduce_cols(int op, int64_t n, int ncols, const double* x,
                    const double* y, double* out, hipStream_t stream) {
  SKT_CHECK(op == SKT_RED_DOT || op == SKT_RED_MAXABS || op == SKT_RED_SUM,
            "bad column reduction op");
  SKT_CHECK(ncols >= 1 && ncols <= 64 && n >= 0 && x && out, "bad argument");
  SKT_CHECK(op != SKT_RED_DOT || y, "y required");
  Scratch* s;
  int st = get_scratch(stream, &s);
  if (st) return st;
  int cp = 1;
  while (cp < ncols) cp <<= 1;
  int rows_per_block = kBlock / cp;
  unsigned nb = grid_for(n, rows_per_block, kMaxBlocks);
  hipLaunchKernelGGL(k_reduce_cols, dim3(nb), dim3(kBlock), 0, stream, op, n,
                     ncols, cp, x, y, s->part);
  hipLaunchKernelGGL(k_reduce_cols_final, dim3(1), dim3(kBlock), 0, stream, op,
                     (int)nb, ncols, s->part, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_row_mean(int64_t n, int k, const double* x, double* out, hipStream_t stream) {
  SKT_CHECK(n >= 0 && k >= 1 && k <= 128 && x && out, "bad argument");
  if (n == 0) return SKT_OK;
  hipLaunchKernelGGL(k_row_mean, dim3(grid_for(n)), dim3(kBlock), 0, stream, n, k, x, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_stats4(int64_t n, const double* x, double* out, hipStream_t stream) {
  SKT_CHECK(n >= 1 && x && out, "statistics of an empty array");
  Scratch* s;
  int st = get_scratch(stream, &s);
  if (st) return st;
  // at most kMaxBlocks chunks of kStatsPart partials fit the scratch
  int64_t chunk = (n + kMaxBlocks - 1) / kMaxBlocks;
  if (chunk < kStatsChunk) chunk = kStatsChunk;
  const int64_t nb = (n + chunk - 1) / chunk;
  SKT_CHECK(nb >= 1 && nb <= kMaxBlocks, "bad statistics grid");
  hipLaunchKernelGGL(k_stats_part, dim3((unsigned)nb), dim3(kBlock), 0, stream, n, chunk, x,
                     s->part);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(kBlock), 0, stream, (int)nb, n,
                     (const double*)s->part, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_percentile_abs(int64_t n, const double* x, double q, double* out,
                       hipStream_t stream) {
  SKT_CHECK(n >= 1 && x && out, "percentile of an empty array");
  SKT_CHECK(q >= 0.0 && q <= 100.0, "Percentiles must be in the range [0, 100]");
  Scratch* s;
  int st = get_scratch(stream, &s);
  if (st) return st;
  return launch_percentile(n, x, q, out, s->part, s->hist, s->sel, stream);
}

int skt_ewise(int op, int64_t n, double a, const double* x, double b, const double* y,
              double* out, hipStream_t stream) {
  SKT_CHECK(op == 0 || op == 1, "op must be 0 (axpby) or 1 (product)");
  SKT_CHECK(n >= 0 && x && out, "bad argument");
  SKT_CHECK(op == 0 || y, "y required for the product");
  if (n == 0) return SKT_OK;
  hipLaunchKernelGGL(k_ewise, dim3(grid_for(n)), dim3(kBlock), 0, stream, op, n, a, x, b, y,
                     out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_oc_normalize(int64_t n, const double* dC, const double* scale,
                     double floor_, double keep, int first, double* running_scale,
                     double* out, hipStream_t stream) {
  SKT_CHECK(n >= 0 && dC && scale && running_scale && out, "bad argument");
  const double omk = 1.0 - keep;
  if (n == 0) {
    hipLaunchKernelGGL(k_store_scalar, dim3(1), dim3(1), 0, stream, scale, floor_, keep, omk,
                       first, running_scale);
  } else {
    // every block derives the same rs; the stored copy is written afterwards
    hipLaunchKernelGGL(k_oc_scale, dim3(grid_for(n)), dim3(kBlock), 0, stream, n, dC, scale,
                       floor_, keep, omk, first, running_scale, out);
    hipLaunchKernelGGL(k_store_scalar, dim3(1), dim3(1), 0, stream, scale, floor_, keep, omk,
                       first, running_scale);
  }
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_oc_kkt(int64_t n, const double* dC_raw, const double* rho, const double* dV,
               const double* bisect_result, const double* running_scale, double rho_min,
               double rho_max, double* out, hipStream_t stream) {
  SKT_CHECK(n >= 1 && dC_raw && rho && dV && bisect_result && running_scale && out,
            "bad argument");
  Scratch* s;
  int st = get_scratch(stream, &s);
  if (st) return st;
  unsigned nb = grid_for(n, kBlock, kMaxBlocks);
  hipLaunchKernelGGL(k_oc_kkt, dim3(nb), dim3(kBlock), 0, stream, n, dC_raw, rho, dV,
                     bisect_result, running_scale, rho_min + 1e-6, rho_max - 1e-6, s->part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, stream, SKT_RED_MAXABS, (int)nb,
                     s->part, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

}  // extern "C"
