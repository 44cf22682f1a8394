// Batched preconditioned CG over ncols right-hand sides that share one
// operator apply per iteration.  Per-column scalars (alpha, beta, r.z, r.r)
// live on the device; dot products and norms are accumulated in float64
// with a fixed reduction order (no float atomics), so results are
// bit-reproducible.  The host only polls the per-column convergence flags
// every few iterations.
//
// Working vectors use a column-blocked layout [ncols/kc][ndof][kc]: with kc
// columns per block the structured operator's workgroups read and write
// contiguous memory.  The caller's F and U stay (ndof, ncols) row-major and
// are converted on entry and exit.  For operators without blocked support
// (tet4) kc = ncols, which is the row-major layout itself.
//
// Preconditioners: Jacobi, or the geometric multigrid V-cycle (mg.hip) on
// structured hex8 grids.
//
// mode SKT_F64: float64 vectors, operator and preconditioner.
// mode SKT_F32: one float32 PCG (float64 dot products) with reliable
// updates: when a column's recurrence residual has dropped by delta^2 in r.r
// since the last update (or claims convergence), the float32 correction is
// folded into the float64 solution, the true residual F - K U is recomputed
// with the float64 operator and replaces the recurrence residual; the search
// direction is kept unless the replacement jumped, so the Krylov space is
// not restarted.  Convergence is only ever declared on a true float64
// residual.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mg.hpp"
#include "operator.hpp"

struct skt_solver {
  skt_operator* op = nullptr;
  int max_cols = 0;
  int64_t n = 0;
  // column-blocked working vectors
  double *x64 = nullptr, *f64 = nullptr, *r64 = nullptr, *p64 = nullptr, *q64 = nullptr,
         *z64 = nullptr, *dinv64 = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *z32 = nullptr, *d32 = nullptr,
        *dinv32 = nullptr;
  double* part = nullptr;   // 4 * kParts * max_cols: area0 | area1
  double* sc = nullptr;     // max_cols * kSc
  double* host = nullptr;   // pinned copy of sc
  // iterations of the last two solves, and set-ups the current coarsest
  // inverse has served beyond its own
  int64_t it_last = 0, it_prev = 0;
  int inv_age = 0;
  // preconditioner
  int precond = SKT_PC_AUTO;
  int used = SKT_PC_JACOBI;  // the one the last solve ran with
  skt::MGConfig mgcfg;
  skt::MGData mg;
  int mg_possible = -1;      // -1 unknown, 0 no, 1 yes
  // optional per-launch timing of the operator apply (hipEvents on the
  // solve stream, read back at the convergence checks); only chunks that
  // start with every column live are timed, since the apply skips the work
  // of stopped columns
  int timing = 0;
  hipEvent_t ev[2 * 32] = {};
  hipEvent_t poll_ev = nullptr;  // completion of a poll's flag copy
  bool timed = false;  // the last chunk's applies were timed
  double apply_ms[2] = {0, 0};
  int64_t n_apply[2] = {0, 0};
  // The PCG iterations between two convergence polls, captured once as a
  // hipGraph per variant (0: plain, 1: reliable update in the first
  // iteration) on a private stream and replayed on the caller's stream.
  // A graph is re-captured when anything baked into its kernel arguments
  // changes: precision, columns, moduli buffers, chunk length, the
  // preconditioner or its buffers.  SKT_GRAPH=0 launches the kernels one by
  // one instead (same kernels, same order).
  struct Chunk {
    hipGraphExec_t exec = nullptr;
    int dtype = -1, k = 0, kc = 0, used = -1, degree = 0;
    int64_t count = 0;
    const void* E = nullptr;
    const double* E64 = nullptr;
    uint64_t mg_gen = 0;
  };
  Chunk chunk[2];
  // The multigrid set-up of a solve (Galerkin products, stencils, coarsest
  // inverse, warm-started power iterations), captured once it is in its
  // steady state (warm start, one-off buffers built) and replayed per solve.
  struct Setup {
    hipGraphExec_t exec = nullptr;
    int dtype = -1, degree = 0, steps = 0;
    const double* E64 = nullptr;
    const void* E = nullptr;
    const void* dinv = nullptr;
    uint64_t mg_gen = 0;
    int lagged = -1;
  };
  Setup setup[2];  // [reuse of the previous coarsest inverse]
  hipStream_t cap = nullptr;
  int graphs = 1;
  int64_t graph_launches = 0, graph_captures = 0;
};

namespace skt {

constexpr int kParts = 512;
constexpr int kSc = 16;
static_assert(kSc == kStopStride, "the column stop flags are the solver's scalar blocks");
constexpr int kCheckJacobi = 32;
constexpr int kCheckMG = 4;

// Iterations between convergence polls with multigrid; SKT_SOLVER_POLL=n
// (1..32) overrides the default.
static int mg_poll_interval() {
  const char* v = std::getenv("SKT_SOLVER_POLL");
  if (v) {
    const int n = std::atoi(v);
    if (n >= 1 && n <= 32) return n;
  }
  return kCheckMG;
}
// scalar slots
enum {
  S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_CONV, S_TOL2, S_BAD,
  S_RREF,    // r.r at the last reliable update (0: no updates, float64 mode)
  S_REQ,     // 1: the column asks for a reliable update
  S_DELTA2,  // reliable-update threshold on r.r / r.r at the last update
  S_NRST,    // float32 breakdowns restarted so far (mixed mode)
  S_RST,     // 1: restart this column's search direction at the update
  S_TRR,     // r.r of the last true (float64) residual
  S_TCONV    // 1: that true residual meets the tolerance
};
static_assert(S_TCONV < 16, "scalar slots");
constexpr int kMaxRestarts = 8;

// Blocked-layout geometry of one solve.
struct Blk {
  int64_t n;      // DOFs
  int k, kc;      // columns, columns per block
  int cp;         // kc rounded up to a power of two (lanes per row)
  int groups;     // k / kc
  unsigned nbx;   // reduction blocks per group
};

static Blk make_blk(int64_t n, int k, int kc) {
  Blk b;
  b.n = n; b.k = k; b.kc = kc;
  b.cp = 1;
  while (b.cp < kc) b.cp <<= 1;
  b.groups = k / kc;
  const int rpb = kBlock / b.cp;
  b.nbx = grid_for(n, rpb, kParts);
  return b;
}

__device__ __forceinline__ bool col_live(const double* sc, int c) {
  return sc[c * kSc + S_CONV] == 0.0;
}

// Thread -> (group g, lane j, first row) of a reduction kernel launched on
// a (nbx, groups) grid.
#define SKT_BLK_SETUP                                              \
  const int j = threadIdx.x % cp;                                  \
  const int rpb = blockDim.x / cp;                                 \
  const int g = blockIdx.y;                                        \
  const int c = g * kc + j;                                        \
  const bool lane_on = j < kc;                                     \
  const int64_t i0 = blockIdx.x * (int64_t)rpb + threadIdx.x / cp; \
  const int64_t istep = (int64_t)gridDim.x * rpb;

// Fixed-order block reduction of per-lane values into part[q][c][bx].
__device__ __forceinline__ void blk_reduce(double* sh, double v, int cp, int kc, int k, int q,
                                           double* part) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s >= cp; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if ((int)threadIdx.x < kc) {
    const int col = blockIdx.y * kc + threadIdx.x;
    part[((int64_t)q * k + col) * gridDim.x + blockIdx.x] = sh[threadIdx.x];
  }
  __syncthreads();
}

// --------------------------------------------------------- layout moves

// dst[(g n + i) kc + j] = src[i k + g kc + j]
__global__ void k_to_blocked(int64_t n, int k, int kc, const double* __restrict__ src,
                             double* __restrict__ dst) {
  const int64_t total = n * k;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / k;
    const int c = (int)(t % k);
    dst[((c / kc) * n + i) * kc + c % kc] = src[t];
  }
}

__global__ void k_from_blocked(int64_t n, int k, int kc, const double* __restrict__ src,
                               double* __restrict__ dst) {
  const int64_t total = n * k;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / k;
    const int c = (int)(t % k);
    dst[t] = src[((c / kc) * n + i) * kc + c % kc];
  }
}

// Dirichlet rows of K_e are identity rows, so U[D] = F[D] (enforce semantics)
__global__ void k_set_fixed(int64_t n, int kc, int64_t total, const uint8_t* __restrict__ fixed,
                            const double* __restrict__ F, double* __restrict__ U) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x)
    if (fixed[(t / kc) % n]) U[t] = F[t];
}

template <typename T>
__global__ void k_dinv(int64_t n, const T* __restrict__ diag, const uint8_t* __restrict__ fixed,
                       T* __restrict__ dinv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dinv[i] = fixed[i] ? T(0) : T(1) / diag[i];
}

// ---------------------------------------------------------- PCG kernels

// r = F - q (F is the enforced right-hand side); partial r.r and F.F
template <typename T>
__global__ void k_resid(int64_t n, int k, int kc, int cp, const double* __restrict__ F,
                        const T* __restrict__ q, T* __restrict__ r, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  SKT_BLK_SETUP
  double rr = 0.0, bb = 0.0;
  if (lane_on) {
    for (int64_t i = i0; i < n; i += istep) {
      const int64_t t = ((int64_t)g * n + i) * kc + j;
      const double f = F[t];
      const T vt = (T)(f - (double)q[t]);
      r[t] = vt;
      rr += (double)vt * (double)vt;
      bb += f * f;
    }
  }
  blk_reduce(sh, rr, cp, kc, k, 0, part);
  blk_reduce(sh, bb, cp, kc, k, 1, part);
}

// partial a.b
template <typename T>
__global__ void k_dot(int64_t n, int k, int kc, int cp, const T* __restrict__ a,
                      const T* __restrict__ b, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  SKT_BLK_SETUP
  double v = 0.0;
  if (lane_on) {
    for (int64_t i = i0; i < n; i += istep) {
      const int64_t t = ((int64_t)g * n + i) * kc + j;
      v += (double)a[t] * (double)b[t];
    }
  }
  blk_reduce(sh, v, cp, kc, k, 0, part);
}

// z = dinv r (when z is null the Jacobi z is formed inline), p = z (when p
// is non-null); partial r.z
template <typename T>
__global__ void k_pinit(int64_t n, int k, int kc, int cp, const T* __restrict__ r,
                        const T* __restrict__ dinv, const T* __restrict__ z, T* __restrict__ p,
                        double* __restrict__ part) {
  __shared__ double sh[kBlock];
  SKT_BLK_SETUP
  double rz = 0.0;
  if (lane_on) {
    for (int64_t i = i0; i < n; i += istep) {
      const int64_t t = ((int64_t)g * n + i) * kc + j;
      const T zv = z ? z[t] : dinv[i] * r[t];
      if (p) p[t] = zv;
      rz += (double)r[t] * (double)zv;
    }
  }
  blk_reduce(sh, rz, cp, kc, k, 0, part);
}

// x += alpha p; r -= alpha q; partial r.z (Jacobi z = dinv r; 0 when dinv
// is null -- the preconditioned product is then taken after the V-cycle)
// and r.r
// (WX: also x += alpha p; otherwise the direction update of the same
// iteration does it, before it overwrites p, which saves a read of p)
template <typename T, bool WX>
__global__ void k_update(int64_t n, int k, int kc, int cp, const double* __restrict__ sc,
                         const T* __restrict__ p, const T* __restrict__ q,
                         const T* __restrict__ dinv, T* __restrict__ x, T* __restrict__ r,
                         double* __restrict__ part) {
  __shared__ double sh[kBlock];
  SKT_BLK_SETUP
  double rz = 0.0, rr = 0.0;
  if (lane_on && col_live(sc, c)) {
    const T alpha = (T)sc[c * kSc + S_ALPHA];
    for (int64_t i = i0; i < n; i += istep) {
      const int64_t t = ((int64_t)g * n + i) * kc + j;
      if constexpr (WX) x[t] += alpha * p[t];
      const T rv = r[t] - alpha * q[t];
      r[t] = rv;
      if (dinv) rz += (double)rv * (double)(dinv[i] * rv);
      rr += (double)rv * (double)rv;
    }
  }
  blk_reduce(sh, rz, cp, kc, k, 0, part);
  blk_reduce(sh, rr, cp, kc, k, 1, part);
}

// p = z + beta p (z null: Jacobi z = dinv r); (nbx, groups) grid as the
// reduction kernels, so the row and column come without 64-bit divisions.
// WX: first x += alpha p with this iteration's alpha (the update k_update<T,
// false> left out), also for a column that stopped at this iteration's
// r.r test: its alpha is nonzero exactly when k_update saw it live.
template <typename T, bool WX>
__global__ void k_pupdate(int64_t n, int k, int kc, int cp, const double* __restrict__ sc,
                          const T* __restrict__ r, const T* __restrict__ dinv,
                          const T* __restrict__ z, T* __restrict__ p, T* __restrict__ x) {
  SKT_BLK_SETUP
  if (!lane_on) return;
  const bool live = col_live(sc, c);
  const T alpha = WX ? (T)sc[c * kSc + S_ALPHA] : T(0);
  const bool upd_x = WX && alpha != T(0);
  if (!live && !upd_x) return;
  const T beta = (T)sc[c * kSc + S_BETA];
  for (int64_t i = i0; i < n; i += istep) {
    const int64_t t = ((int64_t)g * n + i) * kc + j;
    // loads unpredicated (a lane returned above unless it stores something),
    // stores predicated: no load waits inside a branch of its own
    const T pv = p[t];
    const T xv = WX ? x[t] : T(0);
    const T zv = z ? z[t] : dinv[i] * r[t];
    if (upd_x) x[t] = xv + alpha * pv;
    if (live) p[t] = zv + beta * pv;
  }
}

// restart the search direction of columns flagged S_RST: p = z
template <typename T>
__global__ void k_restart_dir(int64_t n, int k, int kc, int cp, const double* __restrict__ sc,
                              const T* __restrict__ r, const T* __restrict__ dinv,
                              const T* __restrict__ z, T* __restrict__ p) {
  SKT_BLK_SETUP
  if (!lane_on || sc[c * kSc + S_RST] == 0.0) return;
  for (int64_t i = i0; i < n; i += istep) {
    const int64_t t = ((int64_t)g * n + i) * kc + j;
    p[t] = z ? z[t] : dinv[i] * r[t];
  }
}

// Sum of the nb partials of quantity q for column c (one wavefront per
// column, fixed order: each lane adds its partials in index order, then the
// wave reduction).  Eight loads are issued before their adds, so a lane
// waits for memory once per eight partials rather than once per partial.
__device__ __forceinline__ double col_sum(const double* part, int k, int nb, int q, int c) {
  const double* p = part + ((int64_t)q * k + c) * nb;
  double a = 0.0;
  int i = threadIdx.x;
  for (; i + 7 * kWave < nb; i += 8 * kWave) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[i + u * kWave];
#pragma unroll
    for (int u = 0; u < 8; ++u) a += v[u];
  }
  for (; i < nb; i += kWave) a += p[i];
  return wave_sum(a);
}

// scalars after initialisation: r.r, F.F from part_rb, r.z from part_rz
__global__ void k_init_final(int nb, int k, const double* __restrict__ part_rb,
                             const double* __restrict__ part_rz, double rtol2, double delta2,
                             double* __restrict__ sc) {
  const int c = blockIdx.x;
  const double rr = col_sum(part_rb, k, nb, 0, c);
  const double bb = col_sum(part_rb, k, nb, 1, c);
  const double rz = col_sum(part_rz, k, nb, 0, c);
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
    s[S_RST] = 0.0;
    if (rr != rr) { s[S_CONV] = 1.0; s[S_BAD] = 1.0; }
  }
}

__global__ void k_alpha(int nb, int k, const double* __restrict__ part, double* __restrict__ sc) {
  const int c = blockIdx.x;
  const double pq = col_sum(part, k, nb, 0, c);
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
}

// r.z from quantity 0 of part_rz (nbz partials per column), r.r from
// quantity 1 of part_rr (nbr partials)
__global__ void k_beta(int nbz, const double* __restrict__ part_rz, int nbr,
                       const double* __restrict__ part_rr, int k, double* __restrict__ sc) {
  const int c = blockIdx.x;
  const double rz = col_sum(part_rz, k, nbz, 0, c);
  const double rr = col_sum(part_rr, k, nbr, 1, c);
  if (threadIdx.x == 0) {
    double* s = sc + c * kSc;
    if (s[S_CONV] == 0.0) {
      s[S_RR] = rr;
      s[S_BETA] = rz / s[S_RZ];
      s[S_RZ] = rz;
      if (rr <= s[S_TOL2]) s[S_CONV] = 1.0;
      if (rr < s[S_DELTA2] * s[S_RREF]) s[S_REQ] = 1.0;
      if (rr != rr || rz != rz) { s[S_CONV] = 1.0; s[S_BAD] = 1.0; }
    }
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

// convergence test on the true residual alone (the float32 recurrence of
// every column claims convergence): S_TRR, S_TCONV
__global__ void k_true_conv(int nb, int k, const double* __restrict__ part_rr,
                            double* __restrict__ sc) {
  const int c = blockIdx.x;
  const double rr = col_sum(part_rr, k, nb, 0, c);
  if (threadIdx.x == 0) {
    double* s = sc + c * kSc;
    s[S_TRR] = rr;
    s[S_TCONV] = (rr <= s[S_TOL2]) ? 1.0 : 0.0;
  }
}

// reliable update inside a mixed-precision iteration: after x += alpha p,
// r -= alpha q the residual has been replaced by the true one (r.r of it in
// part_rr, quantity 0) and preconditioned (r.z in part_rz); this is k_beta
// for that iteration.  The search direction restarts (beta = 0) where the
// true residual left the recurrence behind or grew since the last update
// (as in k_replace_final), where the column had stopped on its recurrence,
// and after a breakdown (up to kMaxRestarts times); convergence is decided
// on the true residual.
__global__ void k_beta_replace(int nbz, const double* __restrict__ part_rz, int nbr,
                               const double* __restrict__ part_rr, int k,
                               double* __restrict__ sc) {
  const int c = blockIdx.x;
  const double rz = col_sum(part_rz, k, nbz, 0, c);
  const double rr = col_sum(part_rr, k, nbr, 0, c);
  if (threadIdx.x == 0) {
    double* s = sc + c * kSc;
    bool rst = rr > 4.0 * s[S_RR] || rr > s[S_RREF] || s[S_CONV] != 0.0;
    if (s[S_BAD] != 0.0 && s[S_NRST] < kMaxRestarts) {
      s[S_BAD] = 0.0;
      s[S_NRST] += 1.0;
      rst = true;
    }
    s[S_BETA] = rst ? 0.0 : rz / s[S_RZ];
    s[S_RZ] = rz;
    s[S_RR] = rr;
    s[S_RREF] = rr;
    s[S_REQ] = 0.0;
    if (s[S_BAD] == 0.0) s[S_CONV] = (rr <= s[S_TOL2]) ? 1.0 : 0.0;
    if (rr != rr || rz != rz) { s[S_CONV] = 1.0; s[S_BAD] = 1.0; }
  }
}

// reliable update, step 3: r.r (float64 residual) and r.z of its float32
// copy replace the recurrence values; convergence is decided here only.
// The search direction is kept unless the true residual has drifted well
// above the recurrence one (the float32 iteration has lost accuracy, and
// the old direction no longer fits the new residual) or the float32
// recurrence broke down; a breakdown is retried up to kMaxRestarts times.
__global__ void k_replace_final(int nb, int k, const double* __restrict__ part_rr,
                                const double* __restrict__ part_rz, double* __restrict__ sc) {
  const int c = blockIdx.x;
  const double rr = col_sum(part_rr, k, nb, 0, c);
  const double rz = col_sum(part_rz, k, nb, 0, c);
  if (threadIdx.x == 0) {
    double* s = sc + c * kSc;
    // restart when the true residual left the recurrence behind, or grew
    // since the previous update (the float32 iteration has diverged)
    s[S_RST] = (rr > 4.0 * s[S_RR] || rr > s[S_RREF]) ? 1.0 : 0.0;
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
}

// ------------------------------------------------------------ host side

// The host waits for the flag copy by querying its event in a loop rather
// than by a blocking synchronisation, whose wake-up left the GPU idle for
// tens of microseconds at every poll.
static int poll(skt_solver* s, int k, hipStream_t st) {
  SKT_HIP(hipMemcpyAsync(s->host, s->sc, sizeof(double) * kSc * k, hipMemcpyDeviceToHost, st));
  SKT_HIP(hipEventRecord(s->poll_ev, st));
  hipError_t e;
  while ((e = hipEventQuery(s->poll_ev)) == hipErrorNotReady) {
  }
  SKT_HIP(e);
  return SKT_OK;
}

static bool all_flag(const skt_solver* s, int k, int slot) {
  for (int c = 0; c < k; ++c)
    if (s->host[c * kSc + slot] == 0.0) return false;
  return true;
}

static bool any_flag(const skt_solver* s, int k, int slot) {
  for (int c = 0; c < k; ++c)
    if (s->host[c * kSc + slot] != 0.0) return true;
  return false;
}

// column block of the solver layout for k columns
static int kc_for(const skt_operator* op, int k) {
  if (op->mesh != SKT_MESH_HEX8 || !op->modal_ok) return k;
  if (op->ncomp == 3) return k % 2 == 0 ? 2 : 1;
  return k % 4 == 0 ? 4 : (k % 2 == 0 ? 2 : 1);
}

// Vectors of one precision for the PCG loop.
template <typename T>
struct Vecs {
  T *x, *r, *p, *q, *z, *dinv;
  const void* E;
  int dtype;
};

// z = M r and the partials of r.z (quantity 0): fused into the V-cycle's
// last smoothing step when they fit (into `fused`, *nb = their count per
// column), else a separate product into `area` (*nb = b.nbx).  Jacobi: z is
// formed inline by the vector kernels and r.z is already in `area`.
template <typename T>
static int precondition(skt_solver* s, const Blk& b, const Vecs<T>& v, double* area,
                        double* fused, const double** rz_part, int* nb, const double* stop,
                        hipStream_t st) {
  *rz_part = area;
  *nb = (int)b.nbx;
  if (s->used == SKT_PC_MG) {
    int slots = 0;
    int stt = mg_vcycle(&s->mg, s->op, v.E, v.dinv, b.k, b.kc, v.r, v.z, v.q, st, fused,
                        2 * kParts, &slots, stop);
    if (stt) return stt;
    if (slots > 0) {
      *rz_part = fused;
      *nb = slots;
    } else {
      hipLaunchKernelGGL(k_dot<T>, dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k, b.kc,
                         b.cp, (const T*)v.r, (const T*)v.z, area);
    }
  }
  return SKT_OK;
}

// Enqueue `count` PCG iterations; stopped columns (S_CONV) are untouched.
// replace_first (mixed precision): the first iteration replaces the
// updated residual by the true one before preconditioning it (a reliable
// update that reuses the iteration's V-cycle).
template <typename T>
static int pcg_iterations(skt_solver* s, const Blk& b, const Vecs<T>& v, int64_t count,
                          bool replace_first, const double* E64, hipStream_t st) {
  const bool mg = s->used == SKT_PC_MG;
  const char* xu = std::getenv("SKT_PCG_X_IN_UPDATE");
  const bool x_in_update = xu && xu[0] == '1' && xu[1] == 0;
  // area0 (s->part) takes the p.q partials and k_update's r.z / r.r; the
  // fused r.z of the V-cycle goes to area1, which is idle between the
  // initialisation and the reliable updates
  double* area1 = s->part + (size_t)2 * kParts * s->max_cols;
  s->timed = s->timing && !any_flag(s, b.k, S_CONV);
  // a modal operator's apply is timed by its own dispatch (the events
  // bracket the kernel, not the queue around it); others by event records
  const bool modal = s->op->mesh == SKT_MESH_HEX8 && s->op->modal_ok;
  for (int64_t it = 0; it < count; ++it) {
    if (s->timed) {
      if (modal) modal_time_next(s->ev[2 * it], s->ev[2 * it + 1]);
      else SKT_HIP(hipEventRecord(s->ev[2 * it], st));
    }
    int pq_nb = 0;
    // stopped columns' q is not used (alpha = 0), so their work may be skipped
    int stt = op_apply_dot(s->op, v.dtype, b.k, b.kc, v.E, v.p, v.q, s->part, 2 * kParts,
                           &pq_nb, st, s->sc + S_CONV);
    if (stt) return stt;
    if (s->timed) {
      if (!modal) {
        SKT_HIP(hipEventRecord(s->ev[2 * it + 1], st));
      } else if (modal_time_clear()) {
        // no modal launch took the events: record them here instead
        SKT_HIP(hipEventRecord(s->ev[2 * it], st));
        SKT_HIP(hipEventRecord(s->ev[2 * it + 1], st));
      }
    }
    if (pq_nb == 0) {
      hipLaunchKernelGGL(k_dot<T>, dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k, b.kc,
                         b.cp, (const T*)v.p, (const T*)v.q, s->part);
      pq_nb = (int)b.nbx;
    }
    hipLaunchKernelGGL(k_alpha, dim3(b.k), dim3(kWave), 0, st, pq_nb, b.k, s->part, s->sc);
    // x += alpha p goes with the direction update, except before a reliable
    // update, which reads x in between (SKT_PCG_X_IN_UPDATE=1: always in
    // k_update)
    const bool x_now = (it == 0 && replace_first) || x_in_update;
    if (x_now)
      hipLaunchKernelGGL((k_update<T, true>), dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k,
                         b.kc, b.cp, s->sc, (const T*)v.p, (const T*)v.q,
                         mg ? (const T*)nullptr : (const T*)v.dinv, v.x, v.r, s->part);
    else
      hipLaunchKernelGGL((k_update<T, false>), dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n,
                         b.k, b.kc, b.cp, s->sc, (const T*)v.p, (const T*)v.q,
                         mg ? (const T*)nullptr : (const T*)v.dinv, v.x, v.r, s->part);
    const double* rz_part = nullptr;
    int rz_nb = 0;
    if (it == 0 && replace_first) {
      // U += d; r = F - K U (float64, r.r partials into area0); d = 0
      const int64_t total = b.n * b.k;
      const unsigned ge = grid_for(total);
      hipLaunchKernelGGL(k_accumulate, dim3(ge), dim3(kBlock), 0, st, total, s->d32, s->x64);
      if ((stt = op_apply(s->op, SKT_F64, b.k, b.kc, E64, s->x64, s->q64, st))) return stt;
      hipLaunchKernelGGL(k_resid<double>, dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k,
                         b.kc, b.cp, (const double*)s->f64, (const double*)s->q64, s->r64,
                         s->part);
      hipLaunchKernelGGL(k_to32, dim3(ge), dim3(kBlock), 0, st, total, s->r64, s->r32, s->d32);
      // r.z partials into area1 whether fused or not (area0 holds r.r);
      // Jacobi forms z = dinv r inline.  No column skips the V-cycle here:
      // a stopped column may resume on its true residual.
      if ((stt = precondition<T>(s, b, v, area1, area1, &rz_part, &rz_nb, nullptr, st)))
        return stt;
      if (!mg)
        hipLaunchKernelGGL(k_pinit<T>, dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k,
                           b.kc, b.cp, (const T*)v.r, (const T*)v.dinv, (const T*)nullptr,
                           (T*)nullptr, area1);
      hipLaunchKernelGGL(k_beta_replace, dim3(b.k), dim3(kWave), 0, st, rz_nb, rz_part,
                         (int)b.nbx, (const double*)s->part, b.k, s->sc);
    } else {
      // stopped columns' z is not used (k_beta and the direction update
      // skip them)
      if ((stt = precondition<T>(s, b, v, s->part, area1, &rz_part, &rz_nb, s->sc + S_CONV,
                                 st)))
        return stt;
      hipLaunchKernelGGL(k_beta, dim3(b.k), dim3(kWave), 0, st, rz_nb, rz_part, (int)b.nbx,
                         (const double*)s->part, b.k, s->sc);
    }
    if (x_now)
      hipLaunchKernelGGL((k_pupdate<T, false>), dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n,
                         b.k, b.kc, b.cp, s->sc, (const T*)v.r, (const T*)v.dinv,
                         mg ? (const T*)v.z : (const T*)nullptr, v.p, v.x);
    else
      hipLaunchKernelGGL((k_pupdate<T, true>), dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n,
                         b.k, b.kc, b.cp, s->sc, (const T*)v.r, (const T*)v.dinv,
                         mg ? (const T*)v.z : (const T*)nullptr, v.p, v.x);
  }
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

// Enqueue `count` PCG iterations on `st`: replay of the chunk's captured
// graph (captured first if missing or stale), or direct launches when
// graphs are off or the applies are being timed.
template <typename T>
static int run_chunk(skt_solver* s, const Blk& b, const Vecs<T>& v, int64_t count,
                     bool replace_first, const double* E64, hipStream_t st) {
  if (!s->graphs || s->timing)
    return pcg_iterations<T>(s, b, v, count, replace_first, E64, st);
  skt_solver::Chunk& g = s->chunk[replace_first ? 1 : 0];
  const bool fresh = g.exec && g.dtype == v.dtype && g.k == b.k && g.kc == b.kc &&
                     g.used == s->used && g.degree == s->mg.degree && g.count == count &&
                     g.E == v.E && g.E64 == E64 && g.mg_gen == s->mg.gen;
  if (!fresh) {
    if (g.exec) {
      // a replay of the old graph may still be queued on the stream
      SKT_HIP(hipStreamSynchronize(st));
      SKT_HIP(hipGraphExecDestroy(g.exec));
      g.exec = nullptr;
    }
    if (!s->cap) SKT_HIP(hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking));
    SKT_HIP(hipStreamBeginCapture(s->cap, hipStreamCaptureModeThreadLocal));
    s->timed = false;
    int stt = pcg_iterations<T>(s, b, v, count, replace_first, E64, s->cap);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s->cap, &graph);
    if (stt) {
      if (graph) (void)hipGraphDestroy(graph);
      return stt;
    }
    SKT_HIP(e);
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SKT_HIP(e);
    g.dtype = v.dtype; g.k = b.k; g.kc = b.kc; g.used = s->used; g.degree = s->mg.degree;
    g.count = count; g.E = v.E; g.E64 = E64; g.mg_gen = s->mg.gen;
    ++s->graph_captures;
  }
  s->timed = false;
  SKT_HIP(hipGraphLaunch(g.exec, st));
  ++s->graph_launches;
  return SKT_OK;
}

// Fold the apply timings of the last chunk into the totals (events of the
// chunk have completed once the poll has synchronised the stream).
static int collect_timing(skt_solver* s, int dtype, int64_t count) {
  if (s->mg.time_fine)
    if (int st = mg_collect_timing(&s->mg)) return st;
  if (!s->timed) return SKT_OK;
  for (int64_t k = 0; k < count; ++k) {
    float ms = 0.f;
    SKT_HIP(hipEventElapsedTime(&ms, s->ev[2 * k], s->ev[2 * k + 1]));
    s->apply_ms[dtype] += ms;
    s->n_apply[dtype] += 1;
  }
  return SKT_OK;
}

// The multigrid set-up on `st`: replay of its captured graph once the
// set-up has reached its steady state, else direct launches.
template <typename T>
static int run_setup_body(skt_solver* s, const double* E64, const Vecs<T>& v, hipStream_t st,
                          bool lagged, bool reuse) {
  MGData& mg = s->mg;
  const bool steady = mg.warm && (s->op->mesh == SKT_MESH_TET4 ||
                                  (mg.Mch && (s->op->n_fixed == 0 || mg.gflag)));
  if (!s->graphs || s->timing || !steady)
    return mg_setup(&mg, s->op, E64, v.E, v.dinv, st);
  skt_solver::Setup& g = s->setup[reuse ? 1 : 0];
  const int steps = mg_warm_steps(&mg);
  const bool fresh = g.exec && g.dtype == v.dtype && g.degree == mg.degree &&
                     g.steps == steps && g.E64 == E64 && g.E == v.E && g.dinv == v.dinv &&
                     g.mg_gen == mg.gen && g.lagged == (int)lagged;
  if (!fresh) {
    if (g.exec) {
      SKT_HIP(hipStreamSynchronize(st));
      SKT_HIP(hipGraphExecDestroy(g.exec));
      g.exec = nullptr;
    }
    if (!s->cap) SKT_HIP(hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking));
    SKT_HIP(hipStreamBeginCapture(s->cap, hipStreamCaptureModeThreadLocal));
    int stt = mg_setup(&mg, s->op, E64, v.E, v.dinv, s->cap);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s->cap, &graph);
    if (stt) {
      if (graph) (void)hipGraphDestroy(graph);
      return stt;
    }
    SKT_HIP(e);
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SKT_HIP(e);
    g.dtype = v.dtype; g.degree = mg.degree; g.steps = steps; g.E64 = E64; g.E = v.E;
    g.dinv = v.dinv; g.mg_gen = mg.gen; g.lagged = (int)lagged;
    ++s->graph_captures;
  }
  SKT_HIP(hipGraphLaunch(g.exec, st));
  ++s->graph_launches;
  return SKT_OK;
}

// The set-up with the lagged coarsest inverse around it: the previous
// background inverse is cast into place first, and the matrix this set-up
// assembled is inverted in the background during the solve.
template <typename T>
static int run_setup(skt_solver* s, const double* E64, const Vecs<T>& v, hipStream_t st) {
  MGData& mg = s->mg;
  const bool lagged = mg_lagged(&mg);
  // keep the previous coarsest inverse while the solves are steady (the
  // last one took no more iterations than the one before it), for at most
  // mg_coarse_refresh set-ups per inverse: while the design moves fast an
  // outdated inverse costs more iterations than its refresh
  const int max_use = mg_coarse_refresh(&mg);
  const bool steady = s->it_last > 0 && s->it_prev > 0 && s->it_last <= s->it_prev;
  const bool reuse = !lagged && mg.have_inv && s->inv_age + 1 < max_use && steady;
  s->inv_age = reuse ? s->inv_age + 1 : 0;
  mg.reuse_inv = reuse ? 1 : 0;
  int stt = mg_lag_pre(&mg, st);
  if (stt) return stt;
  if ((stt = run_setup_body<T>(s, E64, v, st, lagged, reuse))) return stt;
  return lagged ? mg_lag_post(&mg, st) : SKT_OK;
}

// Pick the preconditioner and, for multigrid, build its hierarchy for E.
template <typename T>
static int setup_precond(skt_solver* s, const Blk& b, const double* E64, const Vecs<T>& v,
                         hipStream_t st) {
  s->used = SKT_PC_JACOBI;
  if (s->precond == SKT_PC_JACOBI) return SKT_OK;
  if (s->mg_possible < 0) {
    std::vector<MGLevel> lv;
    const bool grid = (s->op->mesh == SKT_MESH_HEX8 && s->op->modal_ok) ||
                      (s->op->mesh == SKT_MESH_TET4 && s->op->tet_grid);
    s->mg_possible = (grid && mg_plan(s->op, s->mgcfg, &lv)) ? 1 : 0;
  }
  const bool kc_ok = b.kc <= (s->op->ncomp == 3 ? 2 : 4);
  if (!s->mg_possible || !kc_ok) {
    if (s->precond == SKT_PC_MG) {
      set_error("multigrid needs a structured hex8 grid that coarsens below max_coarse DOFs");
      return SKT_EINVAL;
    }
    return SKT_OK;
  }
  int stt = mg_alloc(&s->mg, s->op, s->mgcfg, v.dtype, s->max_cols);
  if (stt) return stt;
  stt = run_setup<T>(s, E64, v, st);
  if (stt) return stt;
  s->used = SKT_PC_MG;
  // the V-cycle's fine-level launches are timed with the applies
  return mg_set_timing(&s->mg, s->timing);
}

template <typename T>
static int solve_t(skt_solver* s, int mode, const Blk& b, const double* E64, const Vecs<T>& v,
                   double rtol, int maxiter, double inner_rtol, int64_t* iters, hipStream_t st) {
  const bool mixed = mode == SKT_F32;
  const bool mg = s->used == SKT_PC_MG;
  const int64_t total = b.n * b.k;
  const unsigned ge = grid_for(total);
  double* area1 = s->part + (size_t)2 * kParts * s->max_cols;
  const int check = mg ? mg_poll_interval() : kCheckJacobi;
  // true residual of the initial guess (float64 operator)
  int stt = op_apply(s->op, SKT_F64, b.k, b.kc, E64, s->x64, s->q64, st);
  if (stt) return stt;
  hipLaunchKernelGGL(k_resid<double>, dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k,
                     b.kc, b.cp, (const double*)s->f64, (const double*)s->q64, s->r64, s->part);
  if (mixed)
    hipLaunchKernelGGL(k_to32, dim3(ge), dim3(kBlock), 0, st, total, s->r64, s->r32, s->d32);
  if (mg) {
    if ((stt = mg_vcycle(&s->mg, s->op, v.E, v.dinv, b.k, b.kc, v.r, v.z, v.q, st))) return stt;
  }
  hipLaunchKernelGGL(k_pinit<T>, dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k, b.kc,
                     b.cp, (const T*)v.r, (const T*)v.dinv, mg ? (const T*)v.z : (const T*)nullptr,
                     v.p, area1);
  // automatic threshold: multigrid converges in few iterations between
  // updates and tolerates 1e-2; Jacobi-PCG runs long enough for float32
  // drift to matter and takes 1e-1
  const double delta = inner_rtol > 0.0 ? inner_rtol : (mg ? 1e-2 : 1e-1);
  const double delta2 = mixed ? delta * delta : 0.0;
  hipLaunchKernelGGL(k_init_final, dim3(b.k), dim3(kWave), 0, st, (int)b.nbx, b.k, s->part, area1,
                     rtol * rtol, delta2, s->sc);
  SKT_LAUNCH_CHECK();
  if ((stt = poll(s, b.k, st))) return stt;
  int64_t it = 0, updates = 0;
  bool replace_next = false;
  bool conv = all_flag(s, b.k, S_CONV) && !any_flag(s, b.k, S_BAD);
  while (!conv && it < maxiter && !any_flag(s, b.k, S_BAD)) {
    const int64_t chunk = std::min<int64_t>(check, maxiter - it);
    if ((stt = run_chunk<T>(s, b, v, chunk, replace_next, E64, st))) return stt;
    if (replace_next) ++updates;
    replace_next = false;
    it += chunk;
    if ((stt = poll(s, b.k, st))) return stt;
    if ((stt = collect_timing(s, v.dtype, chunk))) return stt;
    const bool bad = any_flag(s, b.k, S_BAD);
    const bool all_stop = all_flag(s, b.k, S_CONV);
    if (!mixed) {
      if (bad) break;
      conv = all_stop;
      continue;
    }
    if (!(bad || all_stop || any_flag(s, b.k, S_REQ) || it >= maxiter)) continue;
    if (!(bad || all_stop || it >= maxiter)) {
      // only reliable-update requests: done inside the next iteration
      replace_next = true;
      continue;
    }
    // reliable update of every column
    hipLaunchKernelGGL(k_accumulate, dim3(ge), dim3(kBlock), 0, st, total, s->d32, s->x64);
    if ((stt = op_apply(s->op, SKT_F64, b.k, b.kc, E64, s->x64, s->q64, st))) return stt;
    hipLaunchKernelGGL(k_resid<double>, dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k,
                       b.kc, b.cp, (const double*)s->f64, (const double*)s->q64, s->r64, s->part);
    if (all_stop && !bad) {
      // every recurrence claims convergence: test the true residuals first
      // and stop without preconditioning them when they all pass
      hipLaunchKernelGGL(k_true_conv, dim3(b.k), dim3(kWave), 0, st, (int)b.nbx, b.k, s->part,
                         s->sc);
      SKT_LAUNCH_CHECK();
      if ((stt = poll(s, b.k, st))) return stt;
      if (all_flag(s, b.k, S_TCONV)) {
        ++updates;
        for (int c = 0; c < b.k; ++c) s->host[c * kSc + S_RR] = s->host[c * kSc + S_TRR];
        conv = true;
        break;
      }
    }
    hipLaunchKernelGGL(k_to32, dim3(ge), dim3(kBlock), 0, st, total, s->r64, s->r32, s->d32);
    if (mg) {
      if ((stt = mg_vcycle(&s->mg, s->op, v.E, v.dinv, b.k, b.kc, v.r, v.z, v.q, st))) return stt;
    }
    hipLaunchKernelGGL(k_pinit<T>, dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k, b.kc,
                       b.cp, (const T*)v.r, (const T*)v.dinv,
                       mg ? (const T*)v.z : (const T*)nullptr, (T*)nullptr, area1);
    hipLaunchKernelGGL(k_replace_final, dim3(b.k), dim3(kWave), 0, st, (int)b.nbx, b.k, s->part,
                       area1, s->sc);
    hipLaunchKernelGGL(k_restart_dir<T>, dim3(b.nbx, b.groups), dim3(kBlock), 0, st, b.n, b.k,
                       b.kc, b.cp, s->sc,
                       (const T*)v.r, (const T*)v.dinv, mg ? (const T*)v.z : (const T*)nullptr,
                       v.p);
    SKT_LAUNCH_CHECK();
    ++updates;
    if ((stt = poll(s, b.k, st))) return stt;
    if (any_flag(s, b.k, S_BAD)) break;
    conv = all_flag(s, b.k, S_CONV);
  }
  iters[0] = it;
  iters[1] = updates;
  return conv ? SKT_OK : SKT_ENOCONV;
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
  for (double** p : {&s->x64, &s->f64, &s->r64, &s->p64, &s->q64, &s->z64})
    SKT_HIP(hipMalloc(p, sizeof(double) * nv));
  for (float** p : {&s->r32, &s->p32, &s->q32, &s->z32, &s->d32})
    SKT_HIP(hipMalloc(p, sizeof(float) * nv));
  SKT_HIP(hipMalloc(&s->dinv64, sizeof(double) * s->n));
  SKT_HIP(hipMalloc(&s->dinv32, sizeof(float) * s->n));
  SKT_HIP(hipMalloc(&s->part, sizeof(double) * kParts * 4 * max_cols));
  SKT_HIP(hipMalloc(&s->sc, sizeof(double) * kSc * max_cols));
  SKT_HIP(hipHostMalloc(&s->host, sizeof(double) * kSc * max_cols, hipHostMallocDefault));
  SKT_HIP(hipEventCreateWithFlags(&s->poll_ev, hipEventDisableTiming));
  const char* g = std::getenv("SKT_GRAPH");
  s->graphs = (g && std::atoi(g) == 0) ? 0 : 1;
  *out = s;
  return SKT_OK;
}

int skt_solver_configure(skt_solver* s, int precond, int smooth_degree, int max_coarse) {
  SKT_CHECK(s, "null solver");
  SKT_CHECK(precond == SKT_PC_JACOBI || precond == SKT_PC_MG || precond == SKT_PC_AUTO,
            "precond must be SKT_PC_JACOBI, SKT_PC_MG or SKT_PC_AUTO");
  SKT_CHECK(smooth_degree >= 1 && smooth_degree <= 8, "smooth_degree must be in [1, 8]");
  SKT_CHECK(max_coarse >= 8 && max_coarse <= 6000, "max_coarse must be in [8, 6000]");
  s->precond = precond;
  if (s->mgcfg.degree != smooth_degree || s->mgcfg.max_coarse != max_coarse) {
    mg_free(&s->mg);
    s->mg_possible = -1;
  }
  s->mgcfg.degree = smooth_degree;
  s->mgcfg.max_coarse = max_coarse;
  return SKT_OK;
}

int skt_solver_info(const skt_solver* s, int* precond_used, int* levels) {
  SKT_CHECK(s && precond_used && levels, "null pointer");
  *precond_used = s->used;
  *levels = s->used == SKT_PC_MG ? (int)s->mg.lv.size() : 1;
  return SKT_OK;
}

int skt_solver_set_timing(skt_solver* s, int enable) {
  SKT_CHECK(s, "null solver");
  if (enable && !s->ev[0]) {
    for (int i = 0; i < 2 * 32; ++i) SKT_HIP(hipEventCreate(&s->ev[i]));
  }
  s->timing = enable ? 1 : 0;
  return SKT_OK;
}

int skt_solver_get_vcycle_timing(skt_solver* s, double* ms, int64_t* n, int reset) {
  SKT_CHECK(s && ms && n, "null pointer");
  for (int i = 0; i < 4; ++i) {
    ms[i] = s->mg.tms[i];
    n[i] = s->mg.tn[i];
    if (reset) {
      s->mg.tms[i] = 0.0;
      s->mg.tn[i] = 0;
    }
  }
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

int skt_solver_graph_stats(const skt_solver* s, int64_t* launches, int64_t* captures) {
  SKT_CHECK(s && launches && captures, "null pointer");
  *launches = s->graph_launches;
  *captures = s->graph_captures;
  return SKT_OK;
}

int skt_solver_destroy(skt_solver* s) {
  if (!s) return SKT_OK;
  for (auto& g : s->chunk)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  for (auto& g : s->setup)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (s->cap) (void)hipStreamDestroy(s->cap);
  for (int i = 0; i < 2 * 32; ++i)
    if (s->ev[i]) (void)hipEventDestroy(s->ev[i]);
  if (s->poll_ev) (void)hipEventDestroy(s->poll_ev);
  mg_free(&s->mg);
  for (double* p : {s->x64, s->f64, s->r64, s->p64, s->q64, s->z64, s->dinv64}) (void)hipFree(p);
  for (float* p : {s->r32, s->p32, s->q32, s->z32, s->d32, s->dinv32}) (void)hipFree(p);
  (void)hipFree(s->part); (void)hipFree(s->sc);
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
  SKT_CHECK(mode == SKT_F64 || inner_rtol < 1.0,
            "inner_rtol (reliable-update threshold) must be < 1 (<= 0: automatic)");
  const int64_t n = s->n;
  const Blk b = make_blk(n, ncols, kc_for(s->op, ncols));
  const int64_t total = n * ncols;
  const unsigned ge = grid_for(total);
  iters[0] = 0;
  iters[1] = 0;
  // blocked copies of F and U; U[D] = F[D]
  hipLaunchKernelGGL(k_to_blocked, dim3(ge), dim3(kBlock), 0, st, n, ncols, b.kc, F, s->f64);
  hipLaunchKernelGGL(k_to_blocked, dim3(ge), dim3(kBlock), 0, st, n, ncols, b.kc, (const double*)U,
                     s->x64);
  hipLaunchKernelGGL(k_set_fixed, dim3(ge), dim3(kBlock), 0, st, n, b.kc, total, s->op->fixed,
                     (const double*)s->f64, s->x64);
  SKT_LAUNCH_CHECK();
  int stt;
  if (mode == SKT_F64) {
    stt = op_diag(s->op, SKT_F64, E64, s->q64, st);
    if (stt) return stt;
    hipLaunchKernelGGL(k_dinv<double>, dim3(grid_for(n)), dim3(kBlock), 0, st, n, s->q64,
                       s->op->fixed, s->dinv64);
    Vecs<double> v{s->x64, s->r64, s->p64, s->q64, s->z64, s->dinv64, E64, SKT_F64};
    if ((stt = setup_precond<double>(s, b, E64, v, st))) return stt;
    stt = solve_t<double>(s, mode, b, E64, v, rtol, maxiter, inner_rtol, iters, st);
  } else {
    stt = op_diag(s->op, SKT_F32, E32, s->q32, st);
    if (stt) return stt;
    hipLaunchKernelGGL(k_dinv<float>, dim3(grid_for(n)), dim3(kBlock), 0, st, n, s->q32,
                       s->op->fixed, s->dinv32);
    Vecs<float> v{s->d32, s->r32, s->p32, s->q32, s->z32, s->dinv32, E32, SKT_F32};
    if ((stt = setup_precond<float>(s, b, E64, v, st))) return stt;
    stt = solve_t<float>(s, mode, b, E64, v, rtol, maxiter, inner_rtol, iters, st);
  }
  if (stt != SKT_OK && stt != SKT_ENOCONV) return stt;
  if (stt == SKT_OK) {
    s->it_prev = s->it_last;
    s->it_last = iters[0];
  }
  // solution back to the caller's layout (stream-ordered; the flags and
  // norms below come from the last poll, so the host does not wait for it)
  hipLaunchKernelGGL(k_from_blocked, dim3(ge), dim3(kBlock), 0, st, n, ncols, b.kc,
                     (const double*)s->x64, U);
  SKT_LAUNCH_CHECK();
  bool bad = false;
  for (int c = 0; c < ncols; ++c) {
    const double* h = s->host + c * kSc;
    relres[c] = h[S_BB] > 0 ? std::sqrt(h[S_RR] / h[S_BB]) : 0.0;
    bad = bad || h[S_BAD] != 0.0;
  }
  if (bad) { set_error("CG breakdown (non-positive p.Kp or NaN)"); return SKT_ENOCONV; }
  if (stt == SKT_ENOCONV) { set_error("CG did not reach rtol within maxiter"); return SKT_ENOCONV; }
  return SKT_OK;
}

}  // extern "C"
