// Batched Jacobi-preconditioned CG over ncols right-hand sides that share one
// operator apply per iteration.  Per-column scalars (alpha, beta, r.z, r.r)
// live on the device; dot products and norms are accumulated in float64 with
// a fixed reduction order (no float atomics), so results are bit-reproducible.
// The host only polls the per-column convergence flags every kCheck steps.
//
// mode SKT_F64: float64 vectors and operator.
// mode SKT_F32: iterative refinement -- float64 residual F - K U with the
// float64 operator, float32 inner PCG on the scaled residual (float64 dots),
// U += s * d accumulated in float64.
#include <cmath>
#include <cstring>
#include <vector>

#include "operator.hpp"

struct skt_solver {
  skt_operator* op = nullptr;
  int max_cols = 0;
  int64_t n = 0;
  double *r64 = nullptr, *p64 = nullptr, *q64 = nullptr, *dinv64 = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *d32 = nullptr, *dinv32 = nullptr;
  double* part = nullptr;   // 4 * kParts * max_cols: area0 | area1
  double* sc = nullptr;     // max_cols * kSc
  double* host = nullptr;   // pinned copy of sc
  // optional per-launch timing of the operator apply (hipEvents on the
  // solve stream, read back at the convergence checks)
  int timing = 0;
  hipEvent_t ev[2 * 32] = {};
  double apply_ms[2] = {0, 0};
  int64_t n_apply[2] = {0, 0};
};

namespace skt {

constexpr int kParts = 512;
constexpr int kSc = 16;
constexpr int kCheck = 32;
// scalar slots
enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_CONV, S_TOL2, S_RZN, S_SCALE, S_BAD };

// Column-parallel block layout: thread -> (row slice, column c = tid % cp).
struct ColMap {
  int cp, rows_per_block;
};

__device__ __forceinline__ bool col_live(const double* sc, int c) {
  return sc[c * kSc + S_CONV] == 0.0;
}

// Fixed-order block reduction over the rows of each column (cp | blockDim).
__device__ __forceinline__ void col_block_reduce(double* sh, double v, int cp, int ncols,
                                                 double* out_part, int nb_stride) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s >= cp; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if ((int)threadIdx.x < ncols) out_part[(int64_t)threadIdx.x * nb_stride + blockIdx.x] = sh[threadIdx.x];
  __syncthreads();
}

// Dirichlet rows of K_e are identity rows, so U[D] = F[D] (enforce semantics)
__global__ void k_set_fixed(int64_t n, int ncols, const uint8_t* __restrict__ fixed,
                            const double* __restrict__ F, double* __restrict__ U) {
  const int64_t total = n * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x)
    if (fixed[t / ncols]) U[t] = F[t];
}

template <typename T>
__global__ void k_dinv(int64_t n, const T* __restrict__ diag, const uint8_t* __restrict__ fixed,
                       T* __restrict__ dinv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dinv[i] = fixed[i] ? T(0) : T(1) / diag[i];
}

// r = F - q (F is the enforced right-hand side); partial r.r and F.F
template <typename T>
__global__ void k_resid(int64_t n, int ncols, int cp, const double* __restrict__ F,
                        const T* __restrict__ q, T* __restrict__ r,
                        double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int c = threadIdx.x % cp;
  const int rpb = blockDim.x / cp;
  double rr = 0.0, bb = 0.0;
  if (c < ncols) {
    for (int64_t i = blockIdx.x * (int64_t)rpb + threadIdx.x / cp; i < n;
         i += (int64_t)gridDim.x * rpb) {
      const int64_t t = i * ncols + c;
      const double f = F[t];
      const double v = f - (double)q[t];
      const T vt = (T)v;
      r[t] = vt;
      rr += (double)vt * (double)vt;
      bb += f * f;
    }
  }
  col_block_reduce(sh, rr, cp, ncols, part, gridDim.x);
  col_block_reduce(sh, bb, cp, ncols, part + (int64_t)ncols * gridDim.x, gridDim.x);
}

// p = dinv r; partial r.z
template <typename T>
__global__ void k_pinit(int64_t n, int ncols, int cp, const T* __restrict__ r,
                        const T* __restrict__ dinv, T* __restrict__ p,
                        double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int c = threadIdx.x % cp;
  const int rpb = blockDim.x / cp;
  double rz = 0.0;
  if (c < ncols) {
    for (int64_t i = blockIdx.x * (int64_t)rpb + threadIdx.x / cp; i < n;
         i += (int64_t)gridDim.x * rpb) {
      const int64_t t = i * ncols + c;
      const T z = dinv[i] * r[t];
      p[t] = z;
      rz += (double)r[t] * (double)z;
    }
  }
  col_block_reduce(sh, rz, cp, ncols, part, gridDim.x);
}

// scalars after initialisation: rr, bb from part[0..], rz from part2
__global__ void k_init_final(int nb, int ncols, const double* __restrict__ part_rb,
                             const double* __restrict__ part_rz, double rtol2,
                             int use_rr_as_bb, double* __restrict__ sc) {
  __shared__ double sh[kBlock / kWave];
  for (int c = 0; c < ncols; ++c) {
    double a = 0, b = 0, z = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      a += part_rb[(int64_t)c * nb + i];
      b += part_rb[(int64_t)(ncols + c) * nb + i];
      z += part_rz[(int64_t)c * nb + i];
    }
    const double rr = block_sum(a, sh);
    __syncthreads();
    const double bb = block_sum(b, sh);
    __syncthreads();
    const double rz = block_sum(z, sh);
    if (threadIdx.x == 0) {
      double* s = sc + c * kSc;
      s[S_RR] = rr;
      s[S_BB] = use_rr_as_bb ? rr : bb;
      s[S_RZ] = rz;
      s[S_TOL2] = rtol2 * s[S_BB];
      s[S_CONV] = (rr <= s[S_TOL2]) ? 1.0 : 0.0;
      s[S_ALPHA] = 0.0;
      s[S_BETA] = 0.0;
      s[S_BAD] = 0.0;
    }
    __syncthreads();
  }
}

// partial p.q
template <typename T>
__global__ void k_dot_pq(int64_t n, int ncols, int cp, const T* __restrict__ p,
                         const T* __restrict__ q, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int c = threadIdx.x % cp;
  const int rpb = blockDim.x / cp;
  double v = 0.0;
  if (c < ncols) {
    for (int64_t i = blockIdx.x * (int64_t)rpb + threadIdx.x / cp; i < n;
         i += (int64_t)gridDim.x * rpb) {
      const int64_t t = i * ncols + c;
      v += (double)p[t] * (double)q[t];
    }
  }
  col_block_reduce(sh, v, cp, ncols, part, gridDim.x);
}

__global__ void k_alpha(int nb, int ncols, const double* __restrict__ part,
                        double* __restrict__ sc) {
  __shared__ double sh[kBlock / kWave];
  for (int c = 0; c < ncols; ++c) {
    double a = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) a += part[(int64_t)c * nb + i];
    const double pq = block_sum(a, sh);
    if (threadIdx.x == 0) {
      double* s = sc + c * kSc;
      s[S_PQ] = pq;
      if (s[S_CONV] != 0.0) {
        s[S_ALPHA] = 0.0;
      } else if (!(pq > 0.0)) {  // breakdown (or NaN): stop this column
        s[S_ALPHA] = 0.0;
        s[S_CONV] = 1.0;
        s[S_BAD] = 1.0;
      } else {
        s[S_ALPHA] = s[S_RZ] / pq;
      }
    }
    __syncthreads();
  }
}

// x += alpha p; r -= alpha q; partial r.z (z = dinv r) and r.r
template <typename T>
__global__ void k_update(int64_t n, int ncols, int cp, const double* __restrict__ sc,
                         const T* __restrict__ p, const T* __restrict__ q,
                         const T* __restrict__ dinv, T* __restrict__ x, T* __restrict__ r,
                         double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int c = threadIdx.x % cp;
  const int rpb = blockDim.x / cp;
  double rz = 0.0, rr = 0.0;
  if (c < ncols && col_live(sc, c)) {
    const T alpha = (T)sc[c * kSc + S_ALPHA];
    for (int64_t i = blockIdx.x * (int64_t)rpb + threadIdx.x / cp; i < n;
         i += (int64_t)gridDim.x * rpb) {
      const int64_t t = i * ncols + c;
      x[t] += alpha * p[t];
      const T rv = r[t] - alpha * q[t];
      r[t] = rv;
      const T z = dinv[i] * rv;
      rz += (double)rv * (double)z;
      rr += (double)rv * (double)rv;
    }
  }
  col_block_reduce(sh, rz, cp, ncols, part, gridDim.x);
  col_block_reduce(sh, rr, cp, ncols, part + (int64_t)ncols * gridDim.x, gridDim.x);
}

__global__ void k_beta(int nb, int ncols, const double* __restrict__ part,
                       double* __restrict__ sc) {
  __shared__ double sh[kBlock / kWave];
  for (int c = 0; c < ncols; ++c) {
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      a += part[(int64_t)c * nb + i];
      b += part[(int64_t)(ncols + c) * nb + i];
    }
    const double rz = block_sum(a, sh);
    __syncthreads();
    const double rr = block_sum(b, sh);
    if (threadIdx.x == 0) {
      double* s = sc + c * kSc;
      if (s[S_CONV] == 0.0) {
        s[S_RR] = rr;
        s[S_BETA] = rz / s[S_RZ];
        s[S_RZ] = rz;
        if (rr <= s[S_TOL2]) s[S_CONV] = 1.0;
        if (rr != rr) { s[S_CONV] = 1.0; s[S_BAD] = 1.0; }
      }
    }
    __syncthreads();
  }
}

// p = dinv r + beta p
template <typename T>
__global__ void k_pupdate(int64_t n, int ncols, const double* __restrict__ sc,
                          const T* __restrict__ r, const T* __restrict__ dinv,
                          T* __restrict__ p) {
  const int64_t total = n * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t % ncols);
    if (!col_live(sc, c)) continue;
    const T beta = (T)sc[c * kSc + S_BETA];
    p[t] = dinv[t / ncols] * r[t] + beta * p[t];
  }
}

// r32 = r64 / scale_c (scale = sqrt(r.r)), d32 = 0
__global__ void k_scale_resid(int64_t n, int ncols, const double* __restrict__ r64,
                              const double* __restrict__ scale, float* __restrict__ r32,
                              float* __restrict__ d32) {
  const int64_t total = n * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t % ncols);
    const double s = scale[c * kSc + S_SCALE];
    r32[t] = s > 0.0 ? (float)(r64[t] / s) : 0.0f;
    d32[t] = 0.0f;
  }
}

__global__ void k_set_scale(int ncols, const double* __restrict__ outer, double rtol2,
                            double* __restrict__ inner) {
  if (threadIdx.x != 0) return;
  for (int c = 0; c < ncols; ++c) {
    const double rr = outer[c * kSc + S_RR], bb = outer[c * kSc + S_BB];
    inner[c * kSc + S_SCALE] = (rr > rtol2 * bb && rr > 0.0) ? sqrt(rr) : 0.0;
  }
}

__global__ void k_accumulate(int64_t n, int ncols, const double* __restrict__ scale,
                             const float* __restrict__ d32, double* __restrict__ U) {
  const int64_t total = n * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t % ncols);
    const double s = scale[c * kSc + S_SCALE];
    if (s > 0.0) U[t] += s * (double)d32[t];
  }
}

static int colmap_for(int ncols, int* cp) {
  int c = 1;
  while (c < ncols) c <<= 1;
  *cp = c;
  return kBlock / c;
}

// One PCG run on (x, r, p, q) of type T with preconditioner dinv.  The
// scalars sc must be initialised (k_init_final).  Returns iterations run.
template <typename T>
static int pcg_loop(skt_solver* s, int dtype, int ncols, const void* E, T* x, T* r, T* p,
                    T* q, const T* dinv, int maxiter, int64_t* iters_out, int* all_conv,
                    hipStream_t st) {
  const int64_t n = s->n;
  int cp;
  const int rpb = colmap_for(ncols, &cp);
  const unsigned nb = grid_for(n, rpb, kParts);
  const unsigned ge = grid_for(n * ncols);
  int64_t it = 0;
  *all_conv = 0;
  // initial convergence check
  SKT_HIP(hipMemcpyAsync(s->host, s->sc, sizeof(double) * kSc * ncols, hipMemcpyDeviceToHost, st));
  SKT_HIP(hipStreamSynchronize(st));
  auto done = [&]() {
    for (int c = 0; c < ncols; ++c)
      if (s->host[c * kSc + S_CONV] == 0.0) return false;
    return true;
  };
  if (done()) { *all_conv = 1; *iters_out = 0; return SKT_OK; }
  while (it < maxiter) {
    const int64_t chunk = std::min<int64_t>(kCheck, maxiter - it);
    for (int64_t k = 0; k < chunk; ++k) {
      if (s->timing) SKT_HIP(hipEventRecord(s->ev[2 * k], st));
      int stt = op_apply(s->op, dtype, ncols, E, p, q, st);
      if (stt) return stt;
      if (s->timing) SKT_HIP(hipEventRecord(s->ev[2 * k + 1], st));
      hipLaunchKernelGGL(k_dot_pq<T>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, p, q, s->part);
      hipLaunchKernelGGL(k_alpha, dim3(1), dim3(kBlock), 0, st, (int)nb, ncols, s->part, s->sc);
      hipLaunchKernelGGL(k_update<T>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, s->sc, p, q,
                         dinv, x, r, s->part);
      hipLaunchKernelGGL(k_beta, dim3(1), dim3(kBlock), 0, st, (int)nb, ncols, s->part, s->sc);
      hipLaunchKernelGGL(k_pupdate<T>, dim3(ge), dim3(kBlock), 0, st, n, ncols, s->sc, r, dinv, p);
    }
    SKT_LAUNCH_CHECK();
    it += chunk;
    SKT_HIP(hipMemcpyAsync(s->host, s->sc, sizeof(double) * kSc * ncols, hipMemcpyDeviceToHost, st));
    SKT_HIP(hipStreamSynchronize(st));
    if (s->timing) {
      for (int64_t k = 0; k < chunk; ++k) {
        float ms = 0.f;
        SKT_HIP(hipEventElapsedTime(&ms, s->ev[2 * k], s->ev[2 * k + 1]));
        s->apply_ms[dtype] += ms;
        s->n_apply[dtype] += 1;
      }
    }
    if (done()) { *all_conv = 1; break; }
  }
  *iters_out = it;
  return SKT_OK;
}

}  // namespace skt

using namespace skt;

extern "C" {

int skt_solver_create(skt_operator* op, int max_cols, skt_solver** out) {
  SKT_CHECK(op && out, "null pointer");
  SKT_CHECK(max_cols >= 1 && max_cols <= 64, "max_cols must be in [1, 64]");
  *out = nullptr;
  skt_solver* s = new skt_solver();
  s->op = op;
  s->max_cols = max_cols;
  s->n = op->ndof;
  const size_t nv = (size_t)s->n * max_cols;
  SKT_HIP(hipMalloc(&s->r64, sizeof(double) * nv));
  SKT_HIP(hipMalloc(&s->p64, sizeof(double) * nv));
  SKT_HIP(hipMalloc(&s->q64, sizeof(double) * nv));
  SKT_HIP(hipMalloc(&s->dinv64, sizeof(double) * s->n));
  SKT_HIP(hipMalloc(&s->r32, sizeof(float) * nv));
  SKT_HIP(hipMalloc(&s->p32, sizeof(float) * nv));
  SKT_HIP(hipMalloc(&s->q32, sizeof(float) * nv));
  SKT_HIP(hipMalloc(&s->d32, sizeof(float) * nv));
  SKT_HIP(hipMalloc(&s->dinv32, sizeof(float) * s->n));
  SKT_HIP(hipMalloc(&s->part, sizeof(double) * kParts * 4 * max_cols));
  SKT_HIP(hipMalloc(&s->sc, sizeof(double) * kSc * max_cols * 2));
  SKT_HIP(hipHostMalloc(&s->host, sizeof(double) * kSc * max_cols * 2, hipHostMallocDefault));
  *out = s;
  return SKT_OK;
}

int skt_solver_set_timing(skt_solver* s, int enable) {
  SKT_CHECK(s, "null solver");
  if (enable && !s->ev[0]) {
    for (int i = 0; i < 2 * kCheck; ++i) SKT_HIP(hipEventCreate(&s->ev[i]));
  }
  s->timing = enable ? 1 : 0;
  return SKT_OK;
}

int skt_solver_get_timing(skt_solver* s, double* apply_ms, int64_t* n_apply, int reset) {
  SKT_CHECK(s && apply_ms && n_apply, "null pointer");
  for (int i = 0; i < 2; ++i) {
    apply_ms[i] = s->apply_ms[i];
    n_apply[i] = s->n_apply[i];
    if (reset) { s->apply_ms[i] = 0.0; s->n_apply[i] = 0; }
  }
  return SKT_OK;
}

int skt_solver_destroy(skt_solver* s) {
  if (!s) return SKT_OK;
  for (int i = 0; i < 2 * kCheck; ++i)
    if (s->ev[i]) (void)hipEventDestroy(s->ev[i]);
  (void)hipFree(s->r64); (void)hipFree(s->p64); (void)hipFree(s->q64); (void)hipFree(s->dinv64);
  (void)hipFree(s->r32); (void)hipFree(s->p32); (void)hipFree(s->q32); (void)hipFree(s->d32);
  (void)hipFree(s->dinv32); (void)hipFree(s->part); (void)hipFree(s->sc);
  (void)hipHostFree(s->host);
  delete s;
  return SKT_OK;
}

int skt_solver_solve(skt_solver* s, int mode, int ncols, const double* E64,
                     const float* E32, const double* F, double* U, double rtol,
                     int maxiter, double inner_rtol, int64_t* iters, double* relres,
                     hipStream_t st) {
  SKT_CHECK(s && E64 && F && U && iters && relres, "null pointer");
  SKT_CHECK(ncols >= 1 && ncols <= s->max_cols, "ncols exceeds the solver's max_cols");
  SKT_CHECK(mode == SKT_F64 || mode == SKT_F32, "mode must be SKT_F64 or SKT_F32");
  SKT_CHECK(mode == SKT_F64 || E32, "float32 moduli are required in SKT_F32 mode");
  SKT_CHECK(rtol > 0.0 && maxiter >= 1, "rtol must be > 0 and maxiter >= 1");
  const int64_t n = s->n;
  int cp;
  const int rpb = colmap_for(ncols, &cp);
  const unsigned nb = grid_for(n, rpb, kParts);
  const unsigned ge = grid_for(n * ncols);
  const unsigned gn = grid_for(n);
  double* area1 = s->part + (size_t)2 * kParts * s->max_cols;  // r.z partials
  iters[0] = 0;
  iters[1] = 0;
  hipLaunchKernelGGL(k_set_fixed, dim3(ge), dim3(kBlock), 0, st, n, ncols, s->op->fixed, F,
                     U);
  int stt;
  if (mode == SKT_F64) {
    stt = op_diag(s->op, SKT_F64, E64, s->q64, st);
    if (stt) return stt;
    hipLaunchKernelGGL(k_dinv<double>, dim3(gn), dim3(kBlock), 0, st, n, s->q64,
                       s->op->fixed, s->dinv64);
    stt = op_apply(s->op, SKT_F64, ncols, E64, U, s->q64, st);
    if (stt) return stt;
    hipLaunchKernelGGL(k_resid<double>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, F,
                       s->q64, s->r64, s->part);
    hipLaunchKernelGGL(k_pinit<double>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, s->r64,
                       s->dinv64, s->p64, area1);
    hipLaunchKernelGGL(k_init_final, dim3(1), dim3(kBlock), 0, st, (int)nb, ncols, s->part,
                       area1, rtol * rtol, 0, s->sc);
    SKT_LAUNCH_CHECK();
    int conv = 0;
    stt = pcg_loop<double>(s, SKT_F64, ncols, E64, U, s->r64, s->p64, s->q64, s->dinv64,
                           maxiter, &iters[0], &conv, st);
    if (stt) return stt;
    iters[1] = 1;
    bool bad = false;
    for (int c = 0; c < ncols; ++c) {
      const double* h = s->host + c * kSc;
      relres[c] = h[S_BB] > 0 ? std::sqrt(h[S_RR] / h[S_BB]) : 0.0;
      bad = bad || h[S_BAD] != 0.0;
    }
    if (bad) { set_error("CG breakdown (non-positive p.Kp or NaN)"); return SKT_ENOCONV; }
    if (!conv) { set_error("CG did not converge within maxiter"); return SKT_ENOCONV; }
    return SKT_OK;
  }
  // ---- mixed precision: float64 refinement around a float32 PCG
  stt = op_diag(s->op, SKT_F32, E32, s->q32, st);
  if (stt) return stt;
  hipLaunchKernelGGL(k_dinv<float>, dim3(gn), dim3(kBlock), 0, st, n, s->q32, s->op->fixed,
                     s->dinv32);
  double* sc_outer = s->sc + kSc * s->max_cols;
  const int max_outer = 60;
  int64_t inner_total = 0;
  bool converged = false;
  for (int outer = 0; outer < max_outer; ++outer) {
    stt = op_apply(s->op, SKT_F64, ncols, E64, U, s->q64, st);
    if (stt) return stt;
    hipLaunchKernelGGL(k_resid<double>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, F,
                       s->q64, s->r64, s->part);
    // use the r.r partials also as the "r.z" slot (unused here)
    hipLaunchKernelGGL(k_init_final, dim3(1), dim3(kBlock), 0, st, (int)nb, ncols, s->part,
                       s->part, rtol * rtol, 0, sc_outer);
    SKT_LAUNCH_CHECK();
    SKT_HIP(hipMemcpyAsync(s->host + kSc * s->max_cols, sc_outer, sizeof(double) * kSc * ncols,
                           hipMemcpyDeviceToHost, st));
    SKT_HIP(hipStreamSynchronize(st));
    iters[1] = outer;
    bool all = true;
    for (int c = 0; c < ncols; ++c) {
      const double* h = s->host + kSc * s->max_cols + c * kSc;
      relres[c] = h[S_BB] > 0 ? std::sqrt(h[S_RR] / h[S_BB]) : 0.0;
      if (!(h[S_RR] <= h[S_TOL2])) all = false;
      if (h[S_RR] != h[S_RR]) { set_error("NaN residual in refinement"); return SKT_ENOCONV; }
    }
    if (all) { converged = true; break; }
    if (inner_total >= maxiter) break;
    hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(64), 0, st, ncols, sc_outer, rtol * rtol,
                       s->sc);
    hipLaunchKernelGGL(k_scale_resid, dim3(ge), dim3(kBlock), 0, st, n, ncols, s->r64, s->sc,
                       s->r32, s->d32);
    // inner init: bb = r.r of the scaled residual; p = dinv r
    hipLaunchKernelGGL(k_pinit<float>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, s->r32,
                       s->dinv32, s->p32, area1);
    hipLaunchKernelGGL(k_dot_pq<float>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, s->r32,
                       s->r32, s->part);
    hipLaunchKernelGGL(k_init_final, dim3(1), dim3(kBlock), 0, st, (int)nb, ncols, s->part,
                       area1, inner_rtol * inner_rtol, 1, s->sc);
    // k_init_final overwrote S_SCALE? no: it leaves slot S_SCALE untouched
    SKT_LAUNCH_CHECK();
    int conv = 0;
    int64_t it = 0;
    stt = pcg_loop<float>(s, SKT_F32, ncols, E32, s->d32, s->r32, s->p32, s->q32, s->dinv32,
                          (int)std::max<int64_t>(1, maxiter - inner_total), &it, &conv, st);
    if (stt) return stt;
    inner_total += it;
    hipLaunchKernelGGL(k_accumulate, dim3(ge), dim3(kBlock), 0, st, n, ncols, s->sc, s->d32, U);
    SKT_LAUNCH_CHECK();
  }
  iters[0] = inner_total;
  if (!converged) {
    set_error("mixed-precision refinement did not reach rtol");
    return SKT_ENOCONV;
  }
  return SKT_OK;
}

}  // extern "C"
