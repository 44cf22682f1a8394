// Robin terms of the heat conduction path on structured hex8 grids
// (fea/solver_heat.py:552-627,705-980; mesh/task_heat.py:60-139).
//
// * Nodal densities with the reference's pairing: np.add.at(rho_nodal,
//   t.ravel(), np.repeat(rho, 8)) adds rho[k / 8] at node t.ravel()[k]; for
//   node n the terms come from its corner slots a = 0..7 in that order
//   (k = a n_el + e), divided by the number of elements of the node.
// * The volume ("virtual") Robin forms int h a(rho) |grad rho| u v and
//   int h a(rho) |grad rho| T_env v, a = rho^p (1 - rho)^q, assembled node by
//   node into a 27-point stencil (no atomics): a node's thread visits its
//   <= 8 elements and their quadrature points in a fixed order.
// * The nodal form of the explicit Robin compliance gradient and its map to
//   elements (sum over the 8 corners of nodal / count).
// * An extra 27-point stencil attached to a scalar hex8 operator: op_apply
//   adds it on free rows (operators.hip), so the CG solves with conduction
//   plus Robin while the multigrid preconditioner keeps the conduction part.
#include <cmath>

#include "operator.hpp"

namespace skt {

// corner a of an element: offsets (kCx, kCy, kCz) (VTK hexahedron order)
__constant__ int kCx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ int kCy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__constant__ int kCz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

__device__ __forceinline__ int corner_of(int cx, int cy, int cz) {
  return cz * 4 + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0));
}

struct Gauss {
  int n;
  double x[4];
  double w[4];
};

static Gauss gauss_rule(int n) {
  Gauss g{};
  g.n = n;
  if (n == 1) {
    g.x[0] = 0.0;
    g.w[0] = 2.0;
  } else if (n == 2) {
    const double a = 1.0 / std::sqrt(3.0);
    g.x[0] = -a; g.x[1] = a;
    g.w[0] = 1.0; g.w[1] = 1.0;
  } else if (n == 3) {
    const double a = std::sqrt(0.6);
    g.x[0] = -a; g.x[1] = 0.0; g.x[2] = a;
    g.w[0] = 5.0 / 9.0; g.w[1] = 8.0 / 9.0; g.w[2] = 5.0 / 9.0;
  } else {
    const double s = 2.0 / 7.0 * std::sqrt(6.0 / 5.0);
    const double a = std::sqrt(3.0 / 7.0 - s), b = std::sqrt(3.0 / 7.0 + s);
    const double wa = (18.0 + std::sqrt(30.0)) / 36.0, wb = (18.0 - std::sqrt(30.0)) / 36.0;
    g.x[0] = -b; g.x[1] = -a; g.x[2] = a; g.x[3] = b;
    g.w[0] = wb; g.w[1] = wa; g.w[2] = wa; g.w[3] = wb;
  }
  return g;
}

__global__ void k_nodal_pairing(int64_t nx, int64_t ny, int64_t nz,
                                const double* __restrict__ rho, double* __restrict__ rn,
                                double* __restrict__ cnt) {
  const int64_t nnx = nx + 1, nny = ny + 1, nn = nnx * nny * (nz + 1);
  const int64_t ne = nx * ny * nz;
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = n % nnx, j = (n / nnx) % nny, k = n / (nnx * nny);
    double acc = 0.0, c = 0.0;
    for (int a = 0; a < 8; ++a) {
      const int64_t ex = i - kCx[a], ey = j - kCy[a], ez = k - kCz[a];
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
      const int64_t e = ex + nx * (ey + ny * ez);
      acc += rho[(a * ne + e) / 8];
      c += 1.0;
    }
    rn[n] = acc / (c > 1.0 ? c : 1.0);
    if (cnt) cnt[n] = c;
  }
}

// Element fields at one quadrature point: shape values N, physical
// gradients dN, rho_h, grad rho_h.
struct QP {
  double N[8];
  double dN[8][3];
};

__device__ __forceinline__ void qp_eval(double xi, double eta, double zeta, double sx, double sy,
                                        double sz, QP& q) {
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const double s0 = 2.0 * kCx[a] - 1.0, s1 = 2.0 * kCy[a] - 1.0, s2 = 2.0 * kCz[a] - 1.0;
    const double f0 = 1.0 + s0 * xi, f1 = 1.0 + s1 * eta, f2 = 1.0 + s2 * zeta;
    q.N[a] = f0 * f1 * f2 / 8.0;
    q.dN[a][0] = s0 * f1 * f2 / 8.0 * sx;
    q.dN[a][1] = s1 * f0 * f2 / 8.0 * sy;
    q.dN[a][2] = s2 * f0 * f1 / 8.0 * sz;
  }
}

// One node: its stencil row of the virtual Robin matrix and its load.
__global__ void k_robin_virtual(int64_t nx, int64_t ny, int64_t nz, double hx, double hy,
                                double hz, Gauss g, const double* __restrict__ rn, double p,
                                double qexp, double hcoef, double Tenv,
                                double* __restrict__ st, double* __restrict__ L) {
  const int64_t nnx = nx + 1, nny = ny + 1, nn = nnx * nny * (nz + 1);
  const double detJ = hx * hy * hz / 8.0;
  const double sx = 2.0 / hx, sy = 2.0 / hy, sz = 2.0 / hz;
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = n % nnx, j = (n / nnx) % nny, k = n / (nnx * nny);
    double row[27];
#pragma unroll
    for (int o = 0; o < 27; ++o) row[o] = 0.0;
    double lv = 0.0;
    for (int sl = 0; sl < 8; ++sl) {          // elements of the node, ascending
      const int64_t ex = i - 1 + (sl & 1), ey = j - 1 + ((sl >> 1) & 1), ez = k - 1 + (sl >> 2);
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
      const int an = corner_of((int)(i - ex), (int)(j - ey), (int)(k - ez));
      double r8[8];
#pragma unroll
      for (int b = 0; b < 8; ++b)
        r8[b] = rn[(ex + kCx[b]) + nnx * ((ey + kCy[b]) + nny * (ez + kCz[b]))];
      for (int c = 0; c < g.n; ++c)
        for (int bq = 0; bq < g.n; ++bq)
          for (int aq = 0; aq < g.n; ++aq) {
            QP q;
            qp_eval(g.x[aq], g.x[bq], g.x[c], sx, sy, sz, q);
            const double wd = g.w[aq] * g.w[bq] * g.w[c] * detJ;
            double rh = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
              rh += q.N[b] * r8[b];
              gx += q.dN[b][0] * r8[b];
              gy += q.dN[b][1] * r8[b];
              gz += q.dN[b][2] * r8[b];
            }
            const double gn = sqrt(gx * gx + gy * gy + gz * gz);
            const double cf = hcoef * pow(rh, p) * pow(1.0 - rh, qexp) * gn;
            const double wa = wd * cf * q.N[an];
#pragma unroll
            for (int b = 0; b < 8; ++b) {
              const int o = (kCx[b] - (int)(i - ex) + 1) + 3 * (kCy[b] - (int)(j - ey) + 1) +
                            9 * (kCz[b] - (int)(k - ez) + 1);
              row[o] += wa * q.N[b];
            }
            lv += wa * Tenv;
          }
    }
#pragma unroll
    for (int o = 0; o < 27; ++o) st[(int64_t)o * nn + n] = row[o];
    L[n] = lv;
  }
}

// Heat-exchange functionals per element (fea/solver_heat.py:306-322,380-436):
// den[e] = int_e |grad rho_h| and num[e*ncols + c] = int_e (T_c - T_env)
// |grad rho_h|, rho_h and T_h the trilinear interpolants of nodal values.
__global__ void k_hx_elem(int64_t nx, int64_t ny, int64_t nz, double hx, double hy, double hz,
                          Gauss g, const double* __restrict__ rn, int ncols,
                          const double* __restrict__ T, double Tenv, double* __restrict__ den,
                          double* __restrict__ num) {
  const int64_t nnx = nx + 1, nny = ny + 1, ne = nx * ny * nz;
  const double detJ = hx * hy * hz / 8.0;
  const double sx = 2.0 / hx, sy = 2.0 / hy, sz = 2.0 / hz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    int64_t nd[8];
    double r8[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      nd[b] = (ex + kCx[b]) + nnx * ((ey + kCy[b]) + nny * (ez + kCz[b]));
      r8[b] = rn[nd[b]];
    }
    double dsum = 0.0;
    for (int c = 0; c < ncols; ++c) num[e * ncols + c] = 0.0;
    for (int c3 = 0; c3 < g.n; ++c3)
      for (int bq = 0; bq < g.n; ++bq)
        for (int aq = 0; aq < g.n; ++aq) {
          QP q;
          qp_eval(g.x[aq], g.x[bq], g.x[c3], sx, sy, sz, q);
          const double wd = g.w[aq] * g.w[bq] * g.w[c3] * detJ;
          double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
          for (int b = 0; b < 8; ++b) {
            gx += q.dN[b][0] * r8[b];
            gy += q.dN[b][1] * r8[b];
            gz += q.dN[b][2] * r8[b];
          }
          const double wg = wd * sqrt(gx * gx + gy * gy + gz * gz);
          dsum += wg;
          for (int c = 0; c < ncols; ++c) {
            double th = 0.0;
#pragma unroll
            for (int b = 0; b < 8; ++b) th += q.N[b] * T[nd[b] * ncols + c];
            num[e * ncols + c] += wg * (th - Tenv);
          }
        }
    den[e] = dsum;
  }
}

// Adjoint load of the heat-exchange objective, one node per thread
// (fea/solver_heat.py:393-408,818-832): int -T_env h_eff |grad rho_h| v with
// h_eff the interpolant of the nodal values h rho^p (1 - rho)^q.
__global__ void k_hx_rhs(int64_t nx, int64_t ny, int64_t nz, double hx, double hy, double hz,
                         Gauss g, const double* __restrict__ rn, double p, double qexp,
                         double hcoef, double Tenv, double* __restrict__ L) {
  const int64_t nnx = nx + 1, nny = ny + 1, nn = nnx * nny * (nz + 1);
  const double detJ = hx * hy * hz / 8.0;
  const double sx = 2.0 / hx, sy = 2.0 / hy, sz = 2.0 / hz;
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = n % nnx, j = (n / nnx) % nny, k = n / (nnx * nny);
    double lv = 0.0;
    for (int sl = 0; sl < 8; ++sl) {          // elements of the node, ascending
      const int64_t ex = i - 1 + (sl & 1), ey = j - 1 + ((sl >> 1) & 1), ez = k - 1 + (sl >> 2);
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
      const int an = corner_of((int)(i - ex), (int)(j - ey), (int)(k - ez));
      double r8[8], h8[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        r8[b] = rn[(ex + kCx[b]) + nnx * ((ey + kCy[b]) + nny * (ez + kCz[b]))];
        h8[b] = hcoef * pow(r8[b], p) * pow(1.0 - r8[b], qexp);
      }
      for (int c = 0; c < g.n; ++c)
        for (int bq = 0; bq < g.n; ++bq)
          for (int aq = 0; aq < g.n; ++aq) {
            QP q;
            qp_eval(g.x[aq], g.x[bq], g.x[c], sx, sy, sz, q);
            const double wd = g.w[aq] * g.w[bq] * g.w[c] * detJ;
            double he = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
              he += q.N[b] * h8[b];
              gx += q.dN[b][0] * r8[b];
              gy += q.dN[b][1] * r8[b];
              gz += q.dN[b][2] * r8[b];
            }
            lv += wd * (-Tenv * he * sqrt(gx * gx + gy * gy + gz * gz)) * q.N[an];
          }
    }
    L[n] = lv;
  }
}

// Nodal explicit gradient of one load column c of T (nn x ncols).
__global__ void k_robin_grad_nodal(int64_t nx, int64_t ny, int64_t nz, double hx, double hy,
                                   double hz, Gauss g, const double* __restrict__ rn,
                                   const double* __restrict__ T, int ncols, int col, double p,
                                   double qexp, double hcoef, double Tenv,
                                   double* __restrict__ nodal) {
  const int64_t nnx = nx + 1, nny = ny + 1, nn = nnx * nny * (nz + 1);
  const double detJ = hx * hy * hz / 8.0;
  const double sx = 2.0 / hx, sy = 2.0 / hy, sz = 2.0 / hz;
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = n % nnx, j = (n / nnx) % nny, k = n / (nnx * nny);
    double acc = 0.0;
    for (int sl = 0; sl < 8; ++sl) {
      const int64_t ex = i - 1 + (sl & 1), ey = j - 1 + ((sl >> 1) & 1), ez = k - 1 + (sl >> 2);
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
      const int an = corner_of((int)(i - ex), (int)(j - ey), (int)(k - ez));
      double r8[8], t8[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int64_t m = (ex + kCx[b]) + nnx * ((ey + kCy[b]) + nny * (ez + kCz[b]));
        r8[b] = rn[m];
        t8[b] = T[m * ncols + col];
      }
      for (int c = 0; c < g.n; ++c)
        for (int bq = 0; bq < g.n; ++bq)
          for (int aq = 0; aq < g.n; ++aq) {
            QP q;
            qp_eval(g.x[aq], g.x[bq], g.x[c], sx, sy, sz, q);
            const double wd = g.w[aq] * g.w[bq] * g.w[c] * detJ;
            double rh = 0.0, th = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
              rh += q.N[b] * r8[b];
              th += q.N[b] * t8[b];
              gx += q.dN[b][0] * r8[b];
              gy += q.dN[b][1] * r8[b];
              gz += q.dN[b][2] * r8[b];
            }
            const double gn = sqrt(gx * gx + gy * gy + gz * gz);
            const double gs = gn > 1e-12 ? gn : 1e-12;
            const double a = pow(rh, p) * pow(1.0 - rh, qexp);
            const double da = p * pow(rh, p - 1.0) * pow(1.0 - rh, qexp) -
                              qexp * pow(rh, p) * pow(1.0 - rh, qexp - 1.0);
            const double phi = 2.0 * Tenv * th - th * th;
            const double gv = (gx * q.dN[an][0] + gy * q.dN[an][1] + gz * q.dN[an][2]) / gs;
            acc += wd * hcoef * (da * gn * phi * q.N[an] + a * phi * gv);
          }
    }
    nodal[n] = acc;
  }
}

// out[e * ncols + col] += sum_a nodal[n_a] / max(count[n_a], 1), a = 0..7
__global__ void k_nodal_to_elem_count(int64_t nx, int64_t ny, int64_t nz,
                                      const double* __restrict__ nodal,
                                      const double* __restrict__ cnt, int ncols, int col,
                                      double* __restrict__ out) {
  const int64_t nnx = nx + 1, nny = ny + 1, ne = nx * ny * nz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int64_t m = (ex + kCx[a]) + nnx * ((ey + kCy[a]) + nny * (ez + kCz[a]));
      const double c = cnt[m];
      acc += nodal[m] / (c > 1.0 ? c : 1.0);
    }
    out[e * ncols + col] = acc;
  }
}

// Y (+)= S X for a scalar 27-point stencil (st[o * nn + node]).  Rows of
// fixed DOFs are left alone when a mask is given (identity rows).
// Column c of node m lives at (c / lkc) * gstride + c % lkc + m * dstride.
template <typename T, bool kAdd>
__global__ void k_extra_apply(int64_t nx, int64_t ny, int64_t nz, int ncols, int lkc,
                              int64_t gstride, int64_t dstride, const T* __restrict__ st,
                              const uint8_t* __restrict__ fixed, const T* __restrict__ X,
                              T* __restrict__ Y) {
  const int64_t nnx = nx + 1, nny = ny + 1, nnz = nz + 1, nn = nnx * nny * nnz;
  const int64_t total = nn * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = t / ncols;
    const int c = (int)(t - n * ncols);
    if (fixed && fixed[n]) continue;
    const int64_t base = (int64_t)(c / lkc) * gstride + c % lkc;
    const int64_t i = n % nnx, j = (n / nnx) % nny, k = n / (nnx * nny);
    T acc = T(0);
    for (int o = 0; o < 27; ++o) {
      const int64_t i2 = i + o % 3 - 1, j2 = j + (o / 3) % 3 - 1, k2 = k + o / 9 - 1;
      if (i2 < 0 || i2 >= nnx || j2 < 0 || j2 >= nny || k2 < 0 || k2 >= nnz) continue;
      const int64_t m = i2 + nnx * (j2 + nny * k2);
      acc += st[(int64_t)o * nn + n] * X[base + m * dstride];
    }
    if (kAdd)
      Y[base + n * dstride] += acc;
    else
      Y[base + n * dstride] = acc;
  }
}

__global__ void k_extra_cast(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ void k_extra_diag(int64_t nn, const double* __restrict__ st, double* __restrict__ d64,
                             float* __restrict__ d32) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * blockDim.x) {
    if (d64) d64[n] += st[13 * nn + n];
    if (d32) d32[n] += (float)st[13 * nn + n];
  }
}

int extra_apply_add(const skt_operator* op, int dtype, int ncols, int kc, const void* X, void* Y,
                    hipStream_t s) {
  if (!op->extra64) return SKT_OK;
  const int lkc = kc;
  const int64_t gstride = (kc == ncols) ? 0 : op->ndof * kc;
  const int64_t dstride = (kc == ncols) ? ncols : kc;
  const unsigned g = grid_for(op->n_nodes * ncols);
  if (dtype == SKT_F64)
    hipLaunchKernelGGL((k_extra_apply<double, true>), dim3(g), dim3(kBlock), 0, s, op->nx, op->ny,
                       op->nz, ncols, lkc, gstride, dstride, op->extra64, op->fixed,
                       (const double*)X, (double*)Y);
  else
    hipLaunchKernelGGL((k_extra_apply<float, true>), dim3(g), dim3(kBlock), 0, s, op->nx, op->ny,
                       op->nz, ncols, lkc, gstride, dstride, op->extra32, op->fixed,
                       (const float*)X, (float*)Y);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int extra_diag_add(const skt_operator* op, int dtype, void* diag, hipStream_t s) {
  if (!op->extra64) return SKT_OK;
  hipLaunchKernelGGL(k_extra_diag, dim3(grid_for(op->n_nodes)), dim3(kBlock), 0, s, op->n_nodes,
                     op->extra64, dtype == SKT_F64 ? (double*)diag : nullptr,
                     dtype == SKT_F32 ? (float*)diag : nullptr);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

}  // namespace skt

using namespace skt;

extern "C" {

int skt_hex8_nodal_average_pairing(int64_t nx, int64_t ny, int64_t nz, const double* rho_elem,
                                   double* rho_node, double* count, hipStream_t stream) {
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(rho_elem && rho_node, "null pointer");
  const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL(k_nodal_pairing, dim3(grid_for(nn)), dim3(kBlock), 0, stream, nx, ny, nz,
                     rho_elem, rho_node, count);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_hex8_robin_virtual(int64_t nx, int64_t ny, int64_t nz, double hx, double hy, double hz,
                           int npts, const double* rho_node, double p, double q, double h,
                           double T_env, double* stencil, double* load, hipStream_t stream) {
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(npts >= 1 && npts <= 4, "npts must be in [1, 4]");
  SKT_CHECK(rho_node && stencil && load, "null pointer");
  const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL(k_robin_virtual, dim3(grid_for(nn)), dim3(kBlock), 0, stream, nx, ny, nz,
                     hx, hy, hz, gauss_rule(npts), rho_node, p, q, h, T_env, stencil, load);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_hex8_heat_exchange(int64_t nx, int64_t ny, int64_t nz, double hx, double hy, double hz,
                           int npts, const double* rho_node, int ncols, const double* T,
                           double p, double q, double h, double T_env, double* den, double* num,
                           double* rhs, hipStream_t stream) {
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(npts >= 1 && npts <= 4 && ncols >= 1, "bad npts or ncols");
  SKT_CHECK(rho_node && T && den && num && rhs, "null pointer");
  const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1), ne = nx * ny * nz;
  const Gauss g = gauss_rule(npts);
  hipLaunchKernelGGL(k_hx_elem, dim3(grid_for(ne)), dim3(kBlock), 0, stream, nx, ny, nz, hx, hy,
                     hz, g, rho_node, ncols, T, T_env, den, num);
  hipLaunchKernelGGL(k_hx_rhs, dim3(grid_for(nn)), dim3(kBlock), 0, stream, nx, ny, nz, hx, hy,
                     hz, g, rho_node, p, q, h, T_env, rhs);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_hex8_robin_grad(int64_t nx, int64_t ny, int64_t nz, double hx, double hy, double hz,
                        int npts, const double* rho_node, const double* count, int ncols,
                        const double* T, double p, double q, double h, double T_env,
                        double* work, double* out, hipStream_t stream) {
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(npts >= 1 && npts <= 4 && ncols >= 1, "bad npts or ncols");
  SKT_CHECK(rho_node && count && T && work && out, "null pointer");
  const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1), ne = nx * ny * nz;
  const Gauss g = gauss_rule(npts);
  for (int c = 0; c < ncols; ++c) {
    hipLaunchKernelGGL(k_robin_grad_nodal, dim3(grid_for(nn)), dim3(kBlock), 0, stream, nx, ny,
                       nz, hx, hy, hz, g, rho_node, T, ncols, c, p, q, h, T_env, work);
    hipLaunchKernelGGL(k_nodal_to_elem_count, dim3(grid_for(ne)), dim3(kBlock), 0, stream, nx,
                       ny, nz, (const double*)work, count, ncols, c, out);
  }
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_operator_set_extra(skt_operator* op, const double* stencil, hipStream_t stream) {
  SKT_CHECK(op, "null operator");
  SKT_CHECK(op->mesh == SKT_MESH_HEX8 && op->ncomp == 1,
            "an extra stencil needs a scalar structured hex8 operator");
  const int64_t n = 27 * op->n_nodes;
  if (!stencil) {
    (void)hipFree(op->extra64);
    (void)hipFree(op->extra32);
    op->extra64 = nullptr;
    op->extra32 = nullptr;
    return SKT_OK;
  }
  if (!op->extra64) {
    SKT_HIP(hipMalloc(&op->extra64, sizeof(double) * n));
    SKT_HIP(hipMalloc(&op->extra32, sizeof(float) * n));
  }
  SKT_HIP(hipMemcpyAsync(op->extra64, stencil, sizeof(double) * n, hipMemcpyDeviceToDevice,
                         stream));
  hipLaunchKernelGGL(k_extra_cast, dim3(grid_for(n)), dim3(kBlock), 0, stream, n, op->extra64,
                     op->extra32);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_stencil27_apply(int64_t nx, int64_t ny, int64_t nz, int ncols, const double* stencil,
                        const double* X, double* Y, hipStream_t stream) {
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1 && ncols >= 1, "bad sizes");
  SKT_CHECK(stencil && X && Y && X != Y, "null pointer or in-place apply");
  const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL((k_extra_apply<double, false>), dim3(grid_for(nn * ncols)), dim3(kBlock), 0,
                     stream, nx, ny, nz, ncols, ncols, (int64_t)0, (int64_t)ncols, stencil,
                     (const uint8_t*)nullptr, X, Y);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

}  // extern "C"
