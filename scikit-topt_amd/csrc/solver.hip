// Batched Jacobi-preconditioned CG over ncols right-hand sides that share one
// operator apply per iteration.  Per-column scalars (alpha, beta, r.z, r.r)
// live on the device; dot products and norms are accumulated in float64 with
// a fixed reduction order (no float atomics), so results are bit-reproducible.
// The host only polls the per-column convergence flags every kCheck steps.
//
// mode SKT_F64: float64 vectors and operator.
// mode SKT_F32: one float32 PCG (float64 dot products) with reliable
// updates: when a column's recurrence residual has dropped by delta^2 in r.r
// since the last update (or claims convergence), the float32 correction is
// folded into the float64 solution, the true residual F - K U is recomputed
// with the float64 operator and replaces the recurrence residual; the search
// direction is kept, so the Krylov space is not restarted.  Convergence is
// only ever declared on a true float64 residual.
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
enum {
  S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_CONV, S_TOL2, S_RZN, S_SCALE, S_BAD,
  S_RREF,    // r.r at the last reliable update (0: no updates, float64 mode)
  S_REQ,     // 1: the column asks for a reliable update
  S_DELTA2,  // reliable-update threshold on r.r / r.r at the last update
  S_NRST,    // float32 breakdowns restarted so far (mixed mode)
  S_RST      // 1: restart this column's search direction at the update
};
constexpr int kMaxRestarts = 8;

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

// p = dinv r (skipped when p is null); partial r.z
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
      if (p) p[t] = z;
      rz += (double)r[t] * (double)z;
    }
  }
  col_block_reduce(sh, rz, cp, ncols, part, gridDim.x);
}

// scalars after initialisation: rr, bb from part[0..], rz from part2
__global__ void k_init_final(int nb, int ncols, const double* __restrict__ part_rb,
                             const double* __restrict__ part_rz, double rtol2,
                             double delta2, double* __restrict__ sc) {
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
      s[S_BB] = bb;
      s[S_RZ] = rz;
      s[S_TOL2] = rtol2 * bb;
      s[S_CONV] = (rr <= s[S_TOL2]) ? 1.0 : 0.0;
      s[S_ALPHA] = 0.0;
      s[S_BETA] = 0.0;
      s[S_BAD] = 0.0;
      s[S_RREF] = delta2 > 0.0 ? rr : 0.0;
      s[S_REQ] = 0.0;
      s[S_DELTA2] = delta2;
      s[S_NRST] = 0.0;
      if (rr != rr) { s[S_CONV] = 1.0; s[S_BAD] = 1.0; }
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
        if (rr < s[S_DELTA2] * s[S_RREF]) s[S_REQ] = 1.0;
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

// reliable update, step 1: U += d32 (float64 accumulation)
__global__ void k_accumulate(int64_t total, const float* __restrict__ d32,
                             double* __restrict__ U) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x)
    U[t] += (double)d32[t];
}

// reliable update, step 2: r32 = r64 (the true residual), d32 = 0
__global__ void k_to32(int64_t total, const double* __restrict__ r64, float* __restrict__ r32,
                       float* __restrict__ d32) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    r32[t] = (float)r64[t];
    d32[t] = 0.0f;
  }
}

// reliable update, step 4 (only for columns flagged by step 3): restart the
// search direction, p = dinv r
template <typename T>
__global__ void k_restart_dir(int64_t n, int ncols, const double* __restrict__ sc,
                              const T* __restrict__ r, const T* __restrict__ dinv,
                              T* __restrict__ p) {
  const int64_t total = n * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t % ncols);
    if (sc[c * kSc + S_RST] != 0.0) p[t] = dinv[t / ncols] * r[t];
  }
}

// reliable update, step 3: r.r (float64 residual) and r.z of its float32
// copy replace the recurrence values; convergence is decided here only.
// The search direction is kept unless the true residual has drifted well
// above the recurrence one (the float32 iteration has lost accuracy, and
// the old direction no longer fits the new residual) or the float32
// recurrence broke down; a breakdown is retried up to kMaxRestarts times.
__global__ void k_replace_final(int nb, int ncols, const double* __restrict__ part_rr,
                                const double* __restrict__ part_rz, double* __restrict__ sc) {
  __shared__ double sh[kBlock / kWave];
  for (int c = 0; c < ncols; ++c) {
    double a = 0, z = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      a += part_rr[(int64_t)c * nb + i];
      z += part_rz[(int64_t)c * nb + i];
    }
    const double rr = block_sum(a, sh);
    __syncthreads();
    const double rz = block_sum(z, sh);
    if (threadIdx.x == 0) {
      double* s = sc + c * kSc;
      s[S_RST] = (rr > 4.0 * s[S_RR]) ? 1.0 : 0.0;
      s[S_RR] = rr;
      s[S_RZ] = rz;
      s[S_RREF] = rr;
      s[S_REQ] = 0.0;
      if (s[S_BAD] != 0.0 && s[S_NRST] < kMaxRestarts) {
        s[S_BAD] = 0.0;
        s[S_NRST] += 1.0;
        s[S_RST] = 1.0;
      }
      if (s[S_BAD] == 0.0) s[S_CONV] = (rr <= s[S_TOL2]) ? 1.0 : 0.0;
      if (rr != rr) { s[S_CONV] = 1.0; s[S_BAD] = 1.0; }
    }
    __syncthreads();
  }
}

static int colmap_for(int ncols, int* cp) {
  int c = 1;
  while (c < ncols) c <<= 1;
  *cp = c;
  return kBlock / c;
}

// Copy the per-column scalars to the host and wait for them.
static int poll(skt_solver* s, int ncols, hipStream_t st) {
  SKT_HIP(hipMemcpyAsync(s->host, s->sc, sizeof(double) * kSc * ncols, hipMemcpyDeviceToHost, st));
  SKT_HIP(hipStreamSynchronize(st));
  return SKT_OK;
}

static bool all_flag(const skt_solver* s, int ncols, int slot) {
  for (int c = 0; c < ncols; ++c)
    if (s->host[c * kSc + slot] == 0.0) return false;
  return true;
}

static bool any_flag(const skt_solver* s, int ncols, int slot) {
  for (int c = 0; c < ncols; ++c)
    if (s->host[c * kSc + slot] != 0.0) return true;
  return false;
}

// Enqueue `count` PCG iterations on (x, r, p, q) of type T with the Jacobi
// preconditioner dinv; the scalars sc must be initialised.  Columns that
// have stopped (S_CONV) are left untouched by every kernel.
template <typename T>
static int pcg_iterations(skt_solver* s, int dtype, int ncols, const void* E, T* x, T* r, T* p,
                          T* q, const T* dinv, int64_t count, hipStream_t st) {
  const int64_t n = s->n;
  int cp;
  const int rpb = colmap_for(ncols, &cp);
  const unsigned nb = grid_for(n, rpb, kParts);
  const unsigned ge = grid_for(n * ncols);
  for (int64_t k = 0; k < count; ++k) {
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
  return SKT_OK;
}

// Fold the apply timings of the last chunk into the totals (events of the
// chunk have completed once the poll has synchronised the stream).
static int collect_timing(skt_solver* s, int dtype, int64_t count) {
  if (!s->timing) return SKT_OK;
  for (int64_t k = 0; k < count; ++k) {
    float ms = 0.f;
    SKT_HIP(hipEventElapsedTime(&ms, s->ev[2 * k], s->ev[2 * k + 1]));
    s->apply_ms[dtype] += ms;
    s->n_apply[dtype] += 1;
  }
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
  SKT_CHECK(mode == SKT_F64 || (inner_rtol > 0.0 && inner_rtol < 1.0),
            "inner_rtol (reliable-update threshold) must be in (0, 1)");
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
  const int dt = mode == SKT_F64 ? SKT_F64 : SKT_F32;
  const void* E = mode == SKT_F64 ? (const void*)E64 : (const void*)E32;
  // preconditioner
  int stt = op_diag(s->op, dt, E, mode == SKT_F64 ? (void*)s->q64 : (void*)s->q32, st);
  if (stt) return stt;
  if (mode == SKT_F64)
    hipLaunchKernelGGL(k_dinv<double>, dim3(gn), dim3(kBlock), 0, st, n, s->q64, s->op->fixed,
                       s->dinv64);
  else
    hipLaunchKernelGGL(k_dinv<float>, dim3(gn), dim3(kBlock), 0, st, n, s->q32, s->op->fixed,
                       s->dinv32);
  // true residual of the initial guess
  stt = op_apply(s->op, SKT_F64, ncols, E64, U, s->q64, st);
  if (stt) return stt;
  hipLaunchKernelGGL(k_resid<double>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, F, s->q64,
                     s->r64, s->part);
  if (mode == SKT_F64) {
    hipLaunchKernelGGL(k_pinit<double>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, s->r64,
                       s->dinv64, s->p64, area1);
  } else {
    hipLaunchKernelGGL(k_to32, dim3(ge), dim3(kBlock), 0, st, n * ncols, s->r64, s->r32, s->d32);
    hipLaunchKernelGGL(k_pinit<float>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, s->r32,
                       s->dinv32, s->p32, area1);
  }
  const double delta2 = mode == SKT_F64 ? 0.0 : inner_rtol * inner_rtol;
  hipLaunchKernelGGL(k_init_final, dim3(1), dim3(kBlock), 0, st, (int)nb, ncols, s->part,
                     area1, rtol * rtol, delta2, s->sc);
  SKT_LAUNCH_CHECK();
  if ((stt = poll(s, ncols, st))) return stt;
  int64_t it = 0, updates = 0;
  bool conv = all_flag(s, ncols, S_CONV);
  while (!conv && it < maxiter) {
    const int64_t chunk = std::min<int64_t>(kCheck, maxiter - it);
    if (mode == SKT_F64)
      stt = pcg_iterations<double>(s, SKT_F64, ncols, E64, U, s->r64, s->p64, s->q64,
                                   s->dinv64, chunk, st);
    else
      stt = pcg_iterations<float>(s, SKT_F32, ncols, E32, s->d32, s->r32, s->p32, s->q32,
                                  s->dinv32, chunk, st);
    if (stt) return stt;
    it += chunk;
    if ((stt = poll(s, ncols, st))) return stt;
    if ((stt = collect_timing(s, dt, chunk))) return stt;
    const bool bad = any_flag(s, ncols, S_BAD);
    const bool all_stop = all_flag(s, ncols, S_CONV);
    if (mode == SKT_F64) {
      if (bad) break;
      conv = all_stop;
      continue;
    }
    if (!(bad || all_stop || any_flag(s, ncols, S_REQ) || it >= maxiter)) continue;
    // reliable update of every column
    hipLaunchKernelGGL(k_accumulate, dim3(ge), dim3(kBlock), 0, st, n * ncols, s->d32, U);
    stt = op_apply(s->op, SKT_F64, ncols, E64, U, s->q64, st);
    if (stt) return stt;
    hipLaunchKernelGGL(k_resid<double>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, F,
                       s->q64, s->r64, s->part);
    hipLaunchKernelGGL(k_to32, dim3(ge), dim3(kBlock), 0, st, n * ncols, s->r64, s->r32, s->d32);
    hipLaunchKernelGGL(k_pinit<float>, dim3(nb), dim3(kBlock), 0, st, n, ncols, cp, s->r32,
                       s->dinv32, (float*)nullptr, area1);
    hipLaunchKernelGGL(k_replace_final, dim3(1), dim3(kBlock), 0, st, (int)nb, ncols, s->part,
                       area1, s->sc);
    hipLaunchKernelGGL(k_restart_dir<float>, dim3(ge), dim3(kBlock), 0, st, n, ncols, s->sc,
                       s->r32, s->dinv32, s->p32);
    SKT_LAUNCH_CHECK();
    ++updates;
    if ((stt = poll(s, ncols, st))) return stt;
    if (any_flag(s, ncols, S_BAD)) break;
    conv = all_flag(s, ncols, S_CONV);
  }
  iters[0] = it;
  iters[1] = updates;
  bool bad = false;
  for (int c = 0; c < ncols; ++c) {
    const double* h = s->host + c * kSc;
    relres[c] = h[S_BB] > 0 ? std::sqrt(h[S_RR] / h[S_BB]) : 0.0;
    bad = bad || h[S_BAD] != 0.0;
  }
  if (bad) { set_error("CG breakdown (non-positive p.Kp or NaN)"); return SKT_ENOCONV; }
  if (!conv) { set_error("CG did not reach rtol within maxiter"); return SKT_ENOCONV; }
  return SKT_OK;
}

}  // extern "C"
