// Robin terms of the heat conduction path on tet4 meshes
// (fea/solver_heat.py:306-624,705-980; mesh/task_heat.py:60-139), with P1
// fields:
//
// * nodal densities with the reference's pairing, np.add.at(rho_nodal,
//   t.ravel(), np.repeat(rho, 4)): term k = a n_el + e adds rho[k / 4] at
//   node t[a][e], in ascending k (corner slot a = 0..3, then element);
// * the volume ("virtual") Robin forms int h a(rho) |grad rho| u v and
//   int h a(rho) |grad rho| T_env v, a = rho^p (1 - rho)^q, as 4 x 4 element
//   matrices (packed 00 11 22 33 01 02 03 12 13 23) and element loads, on a
//   tetrahedral rule: order 1 the centroid, order 2 the symmetric 4-point rule
//   of degree 2 (one barycentric coordinate (5 + 3 sqrt 5) / 20, the others
//   (5 - sqrt 5) / 20, weights 1/4), int_e f = V sum_q w_q f(lambda_q);
// * the heat-exchange functionals and adjoint load (exact: |grad rho_h| is
//   constant on a tet and T_h, h_eff linear);
// * the nodal explicit Robin gradient and its map to elements;
// * x_e^T R_e x_e of the extra element matrices (the Robin part of T^T K T).
//
// Nodal sums run over the node -> element lists in their ascending element
// order (no atomics).  The extra element matrices ride on the scalar tet
// operator (skt_tet4_set_extra): every apply adds R_e x_e on free rows and
// the diagonal includes R_aa, while a multigrid hierarchy built on the grid
// of cells keeps the conduction part only.
#include <cmath>

#include "operator.hpp"

namespace skt {

struct TetRule {
  int n;
  double lam[4][4];  // [point][corner]
  double w[4];
};

static TetRule tet_rule(int order) {
  TetRule r{};
  if (order <= 1) {
    r.n = 1;
    for (int a = 0; a < 4; ++a) r.lam[0][a] = 0.25;
    r.w[0] = 1.0;
    return r;
  }
  const double A = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0, B = (5.0 - std::sqrt(5.0)) / 20.0;
  r.n = 4;
  for (int q = 0; q < 4; ++q) {
    for (int a = 0; a < 4; ++a) r.lam[q][a] = q == a ? A : B;
    r.w[q] = 0.25;
  }
  return r;
}

__device__ __forceinline__ int tsym(int a, int b) {
  if (a == b) return a;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo == 0 ? 3 + hi : (lo == 1 ? 5 + hi : 9);
}

// rho_h's gradient on element e (constant), from the barycentric gradients
// g[(3 b + d) n_el + e] and the corner values r4
__device__ __forceinline__ void tet_grad(int64_t n_el, int64_t e, const double* __restrict__ g,
                                         const double r4[4], double gr[3]) {
  gr[0] = gr[1] = gr[2] = 0.0;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int d = 0; d < 3; ++d) gr[d] += g[(3 * b + d) * n_el + e] * r4[b];
}

__global__ void k_tet_nodal_pairing(int64_t n_nodes, int64_t n_el,
                                    const int64_t* __restrict__ nptr,
                                    const int32_t* __restrict__ nel,
                                    const uint8_t* __restrict__ nloc,
                                    const double* __restrict__ rho, double* __restrict__ rn,
                                    double* __restrict__ cnt) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < n_nodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q0 = nptr[n], q1 = nptr[n + 1];
    double acc = 0.0;
    // ascending k = a n_el + e: corner slot first, then the (ascending)
    // elements of the list with that slot
    for (int a = 0; a < 4; ++a)
      for (int64_t q = q0; q < q1; ++q)
        if (nloc[q] == a) acc += rho[((int64_t)a * n_el + nel[q]) / 4];
    const double c = (double)(q1 - q0);
    rn[n] = acc / (c > 1.0 ? c : 1.0);
    if (cnt) cnt[n] = c;
  }
}

__global__ void k_tet_robin_virtual(int64_t n_el, TetRule r, const int32_t* __restrict__ conn,
                                    const double* __restrict__ g, const double* __restrict__ vol,
                                    const double* __restrict__ rn, double p, double qexp,
                                    double hcoef, double Tenv, double* __restrict__ R,
                                    double* __restrict__ Le) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_el;
       e += (int64_t)gridDim.x * blockDim.x) {
    double r4[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) r4[b] = rn[conn[b * n_el + e]];
    double gr[3];
    tet_grad(n_el, e, g, r4, gr);
    const double gn = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    const double V = vol[e];
    double m[10], l[4];
#pragma unroll
    for (int k = 0; k < 10; ++k) m[k] = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) l[a] = 0.0;
    for (int q = 0; q < r.n; ++q) {
      double rh = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) rh += r.lam[q][b] * r4[b];
      const double cf = V * r.w[q] * hcoef * pow(rh, p) * pow(1.0 - rh, qexp) * gn;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        l[a] += cf * Tenv * r.lam[q][a];
#pragma unroll
        for (int b = a; b < 4; ++b) m[tsym(a, b)] += cf * r.lam[q][a] * r.lam[q][b];
      }
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) R[(int64_t)k * n_el + e] = m[k];
#pragma unroll
    for (int a = 0; a < 4; ++a) Le[(int64_t)a * n_el + e] = l[a];
  }
}

// out[n] = sum of C[a n_el + e] over the node's (element, corner) pairs, in
// the list's ascending element order
__global__ void k_tet_corner_gather(int64_t n_nodes, int64_t n_el,
                                    const int64_t* __restrict__ nptr,
                                    const int32_t* __restrict__ nel,
                                    const uint8_t* __restrict__ nloc,
                                    const double* __restrict__ C, double* __restrict__ out) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < n_nodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int64_t q = nptr[n]; q < nptr[n + 1]; ++q)
      acc += C[(int64_t)nloc[q] * n_el + nel[q]];
    out[n] = acc;
  }
}

// Heat-exchange functionals per element: den[e] = V |grad rho_h| and
// num[e ncols + c] = V |grad rho_h| (mean of the corner T_c - T_env), exact
// for P1 fields.  Adjoint load per element corner:
// Le[a n_el + e] = -T_env |grad rho_h| sum_b h_b V (1 + d_ab) / 20, with h_b
// the nodal h rho^p (1 - rho)^q.
__global__ void k_tet_hx(int64_t n_el, const int32_t* __restrict__ conn,
                         const double* __restrict__ g, const double* __restrict__ vol,
                         const double* __restrict__ rn, int ncols, const double* __restrict__ T,
                         double p, double qexp, double hcoef, double Tenv,
                         double* __restrict__ den, double* __restrict__ num,
                         double* __restrict__ Le) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_el;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t nd[4];
    double r4[4], h4[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      nd[b] = conn[b * n_el + e];
      r4[b] = rn[nd[b]];
      h4[b] = hcoef * pow(r4[b], p) * pow(1.0 - r4[b], qexp);
    }
    double gr[3];
    tet_grad(n_el, e, g, r4, gr);
    const double V = vol[e];
    const double vg = V * sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    den[e] = vg;
    for (int c = 0; c < ncols; ++c) {
      double tm = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) tm += T[nd[b] * ncols + c];
      num[e * ncols + c] = vg * (0.25 * tm - Tenv);
    }
    const double hs = h4[0] + h4[1] + h4[2] + h4[3];
#pragma unroll
    for (int a = 0; a < 4; ++a) Le[(int64_t)a * n_el + e] = -Tenv * vg * (hs + h4[a]) / 20.0;
  }
}

// Element corners of the nodal explicit Robin gradient of load column col:
// int h (da |grad rho| phi N_a + a phi (grad rho / |grad rho|) . grad N_a),
// phi = 2 T_env T - T^2, |grad rho| floored at 1e-12 in the division.
__global__ void k_tet_robin_grad(int64_t n_el, TetRule r, const int32_t* __restrict__ conn,
                                 const double* __restrict__ g, const double* __restrict__ vol,
                                 const double* __restrict__ rn, const double* __restrict__ T,
                                 int ncols, int col, double p, double qexp, double hcoef,
                                 double Tenv, double* __restrict__ Ce) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_el;
       e += (int64_t)gridDim.x * blockDim.x) {
    double r4[4], t4[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t m = conn[b * n_el + e];
      r4[b] = rn[m];
      t4[b] = T[m * ncols + col];
    }
    double gr[3];
    tet_grad(n_el, e, g, r4, gr);
    const double gn = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    const double gs = gn > 1e-12 ? gn : 1e-12;
    double gv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
      gv[a] = (gr[0] * g[(3 * a + 0) * n_el + e] + gr[1] * g[(3 * a + 1) * n_el + e] +
               gr[2] * g[(3 * a + 2) * n_el + e]) / gs;
    const double V = vol[e];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < r.n; ++q) {
      double rh = 0.0, th = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        rh += r.lam[q][b] * r4[b];
        th += r.lam[q][b] * t4[b];
      }
      const double a = pow(rh, p) * pow(1.0 - rh, qexp);
      const double da = p * pow(rh, p - 1.0) * pow(1.0 - rh, qexp) -
                        qexp * pow(rh, p) * pow(1.0 - rh, qexp - 1.0);
      const double phi = 2.0 * Tenv * th - th * th;
      const double wd = V * r.w[q] * hcoef;
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[k] += wd * (da * gn * phi * r.lam[q][k] + a * phi * gv[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) Ce[(int64_t)k * n_el + e] = acc[k];
  }
}

// out[e ncols + col] = sum_a nodal[t_a] / max(count[t_a], 1), a = 0..3
__global__ void k_tet_nodal_to_elem_count(int64_t n_el, const int32_t* __restrict__ conn,
                                          const double* __restrict__ nodal,
                                          const double* __restrict__ cnt, int ncols, int col,
                                          double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_el;
       e += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t m = conn[a * n_el + e];
      const double c = cnt[m];
      acc += nodal[m] / (c > 1.0 ? c : 1.0);
    }
    out[e * ncols + col] = acc;
  }
}

// out[e ncols + c] = x_e^T R_e x_e
__global__ void k_tet_extra_quad(int64_t n_el, int ncols, const int32_t* __restrict__ conn,
                                 const double* __restrict__ R, const double* __restrict__ X,
                                 double* __restrict__ out) {
  const int64_t total = n_el * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / ncols;
    const int c = (int)(t - e * ncols);
    double x[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) x[b] = X[(int64_t)conn[b * n_el + e] * ncols + c];
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      double ra = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) ra += R[(int64_t)tsym(a, b) * n_el + e] * x[b];
      s += x[a] * ra;
    }
    out[t] = s;
  }
}

__global__ void k_tet_extra_cast(int64_t n, const double* __restrict__ a,
                                 float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

static int check_scalar_tet(const skt_operator* op) {
  SKT_CHECK(op && op->mesh == SKT_MESH_TET4 && op->ncomp == 1,
            "a scalar tet4 operator is required");
  return SKT_OK;
}

}  // namespace skt

using namespace skt;

extern "C" {

int skt_tet4_nodal_average_pairing(const skt_operator* op, const double* rho_elem,
                                   double* rho_node, double* count, hipStream_t stream) {
  SKT_CHECK(op && op->mesh == SKT_MESH_TET4, "a tet4 operator is required");
  SKT_CHECK(rho_elem && rho_node, "null pointer");
  hipLaunchKernelGGL(k_tet_nodal_pairing, dim3(grid_for(op->n_nodes)), dim3(kBlock), 0, stream,
                     op->n_nodes, op->n_el, op->nptr, op->nel, op->nloc, rho_elem, rho_node,
                     count);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_tet4_robin_virtual(const skt_operator* op, int order, const double* rho_node, double p,
                           double q, double h, double T_env, double* elem_mat, double* load,
                           double* work, hipStream_t stream) {
  if (int st = check_scalar_tet(op)) return st;
  SKT_CHECK(order == 1 || order == 2, "order must be 1 or 2");
  SKT_CHECK(rho_node && elem_mat && load && work, "null pointer");
  hipLaunchKernelGGL(k_tet_robin_virtual, dim3(grid_for(op->n_el)), dim3(kBlock), 0, stream,
                     op->n_el, tet_rule(order), op->conn, op->g64, op->vol64, rho_node, p, q, h,
                     T_env, elem_mat, work);
  hipLaunchKernelGGL(k_tet_corner_gather, dim3(grid_for(op->n_nodes)), dim3(kBlock), 0, stream,
                     op->n_nodes, op->n_el, op->nptr, op->nel, op->nloc, (const double*)work,
                     load);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_tet4_heat_exchange(const skt_operator* op, const double* rho_node, int ncols,
                           const double* T, double p, double q, double h, double T_env,
                           double* den, double* num, double* rhs, double* work,
                           hipStream_t stream) {
  if (int st = check_scalar_tet(op)) return st;
  SKT_CHECK(ncols >= 1, "ncols must be >= 1");
  SKT_CHECK(rho_node && T && den && num && rhs && work, "null pointer");
  hipLaunchKernelGGL(k_tet_hx, dim3(grid_for(op->n_el)), dim3(kBlock), 0, stream, op->n_el,
                     op->conn, op->g64, op->vol64, rho_node, ncols, T, p, q, h, T_env, den, num,
                     work);
  hipLaunchKernelGGL(k_tet_corner_gather, dim3(grid_for(op->n_nodes)), dim3(kBlock), 0, stream,
                     op->n_nodes, op->n_el, op->nptr, op->nel, op->nloc, (const double*)work,
                     rhs);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_tet4_robin_grad(const skt_operator* op, int order, const double* rho_node,
                        const double* count, int ncols, const double* T, double p, double q,
                        double h, double T_env, double* work, double* out, hipStream_t stream) {
  if (int st = check_scalar_tet(op)) return st;
  SKT_CHECK(order == 1 || order == 2, "order must be 1 or 2");
  SKT_CHECK(ncols >= 1, "ncols must be >= 1");
  SKT_CHECK(rho_node && count && T && work && out, "null pointer");
  double* corners = work;                       // 4 n_el
  double* nodal = work + 4 * op->n_el;          // n_nodes
  const TetRule r = tet_rule(order);
  for (int c = 0; c < ncols; ++c) {
    hipLaunchKernelGGL(k_tet_robin_grad, dim3(grid_for(op->n_el)), dim3(kBlock), 0, stream,
                       op->n_el, r, op->conn, op->g64, op->vol64, rho_node, T, ncols, c, p, q, h,
                       T_env, corners);
    hipLaunchKernelGGL(k_tet_corner_gather, dim3(grid_for(op->n_nodes)), dim3(kBlock), 0, stream,
                       op->n_nodes, op->n_el, op->nptr, op->nel, op->nloc,
                       (const double*)corners, nodal);
    hipLaunchKernelGGL(k_tet_nodal_to_elem_count, dim3(grid_for(op->n_el)), dim3(kBlock), 0,
                       stream, op->n_el, op->conn, (const double*)nodal, count, ncols, c, out);
  }
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_tet4_set_extra(skt_operator* op, const double* elem_mat, hipStream_t stream) {
  if (int st = check_scalar_tet(op)) return st;
  const int64_t n = 10 * op->n_el;
  if (!elem_mat) {
    (void)hipFree(op->tet_x64);
    (void)hipFree(op->tet_x32);
    op->tet_x64 = nullptr;
    op->tet_x32 = nullptr;
    return SKT_OK;
  }
  if (!op->tet_x64) {
    SKT_HIP(hipMalloc(&op->tet_x64, sizeof(double) * n));
    SKT_HIP(hipMalloc(&op->tet_x32, sizeof(float) * n));
  }
  SKT_HIP(hipMemcpyAsync(op->tet_x64, elem_mat, sizeof(double) * n, hipMemcpyDeviceToDevice,
                         stream));
  hipLaunchKernelGGL(k_tet_extra_cast, dim3(grid_for(n)), dim3(kBlock), 0, stream, n,
                     op->tet_x64, op->tet_x32);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_tet4_extra_quadform(const skt_operator* op, int ncols, const double* X, double* out,
                            hipStream_t stream) {
  if (int st = check_scalar_tet(op)) return st;
  SKT_CHECK(op->tet_x64, "no extra element matrices are set");
  SKT_CHECK(ncols >= 1 && X && out, "bad arguments");
  hipLaunchKernelGGL(k_tet_extra_quad, dim3(grid_for(op->n_el * ncols)), dim3(kBlock), 0,
                     stream, op->n_el, ncols, op->conn, (const double*)op->tet_x64, X, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

}  // extern "C"
