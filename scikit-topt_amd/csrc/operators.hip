// Matrix-free SIMP/RAMP-weighted operators K(E) = sum_e E_e K_e^0 with
// Dirichlet rows replaced by identity, for structured hex8 grids and
// unstructured tet4 meshes, elasticity (3 dof/node) and conduction (1).
//
// Block vectors are (ndof, ncols) row-major: the ncols load cases of one DOF
// are contiguous, and consecutive lanes of a wavefront take consecutive
// (dof, column) pairs so every gather is a coalesced run.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "operator.hpp"

namespace skt {

// Corner order of the local hex nodes (bit 0: x, bit 1: y, bit 2: z is NOT
// used; explicit table instead).
__host__ __device__ constexpr int kCx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__host__ __device__ constexpr int kCy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__host__ __device__ constexpr int kCz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

// local index of the corner (cx, cy, cz)
__host__ __device__ constexpr int corner_id(int cx, int cy, int cz) {
  return cz * 4 + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0));
}

// ------------------------------------------------------------ host K_e^0

static void gauss_rule(int order, std::vector<double>& x, std::vector<double>& w) {
  switch (order) {
    case 1: x = {0.0}; w = {2.0}; break;
    case 2: { double a = 1.0 / std::sqrt(3.0); x = {-a, a}; w = {1.0, 1.0}; } break;
    case 3: { double a = std::sqrt(0.6); x = {-a, 0.0, a}; w = {5.0 / 9, 8.0 / 9, 5.0 / 9}; } break;
    default: {
      double a = std::sqrt(3.0 / 7 - 2.0 / 7 * std::sqrt(6.0 / 5));
      double b = std::sqrt(3.0 / 7 + 2.0 / 7 * std::sqrt(6.0 / 5));
      double wa = (18.0 + std::sqrt(30.0)) / 36.0, wb = (18.0 - std::sqrt(30.0)) / 36.0;
      x = {-b, -a, a, b}; w = {wb, wa, wa, wb};
    } break;
  }
}

static void hex_grads(double xi, double et, double ze, const double h[3], double g[8][3]) {
  for (int a = 0; a < 8; ++a) {
    double sx = 2.0 * kCx[a] - 1, sy = 2.0 * kCy[a] - 1, sz = 2.0 * kCz[a] - 1;
    g[a][0] = sx * (1 + sy * et) * (1 + sz * ze) / 8.0 * (2.0 / h[0]);
    g[a][1] = sy * (1 + sx * xi) * (1 + sz * ze) / 8.0 * (2.0 / h[1]);
    g[a][2] = sz * (1 + sx * xi) * (1 + sy * et) / 8.0 * (2.0 / h[2]);
  }
}

static void hex_K0(int physics, double nu, const double h[3], int order, double* K) {
  const int nd = physics == SKT_ELASTIC ? 24 : 8;
  std::memset(K, 0, sizeof(double) * nd * nd);
  std::vector<double> x, w;
  gauss_rule(order, x, w);
  const double detJ = h[0] * h[1] * h[2] / 8.0;
  const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = 1.0 / (2.0 * (1.0 + nu));
  for (size_t i = 0; i < x.size(); ++i)
    for (size_t j = 0; j < x.size(); ++j)
      for (size_t k = 0; k < x.size(); ++k) {
        double g[8][3];
        hex_grads(x[i], x[j], x[k], h, g);
        const double wq = w[i] * w[j] * w[k] * detJ;
        if (physics == SKT_HEAT) {
          for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b)
              K[a * 8 + b] += wq * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
          continue;
        }
        // B (6 x 24), Voigt [xx, yy, zz, yz, xz, xy] with engineering shear
        double B[6][24];
        std::memset(B, 0, sizeof(B));
        for (int a = 0; a < 8; ++a) {
          const double gx = g[a][0], gy = g[a][1], gz = g[a][2];
          B[0][3 * a] = gx; B[1][3 * a + 1] = gy; B[2][3 * a + 2] = gz;
          B[3][3 * a + 1] = gz; B[3][3 * a + 2] = gy;
          B[4][3 * a + 0] = gz; B[4][3 * a + 2] = gx;
          B[5][3 * a + 0] = gy; B[5][3 * a + 1] = gx;
        }
        double D[6][6];
        std::memset(D, 0, sizeof(D));
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) D[r][c] = lam + (r == c ? 2.0 * mu : 0.0);
        D[3][3] = D[4][4] = D[5][5] = mu;
        double DB[6][24];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 24; ++c) {
            double s = 0;
            for (int t = 0; t < 6; ++t) s += D[r][t] * B[t][c];
            DB[r][c] = s;
          }
        for (int r = 0; r < 24; ++r)
          for (int c = 0; c < 24; ++c) {
            double s = 0;
            for (int t = 0; t < 6; ++t) s += B[t][r] * DB[t][c];
            K[r * 24 + c] += wq * s;
          }
      }
  for (int r = 0; r < nd; ++r)
    for (int c = r + 1; c < nd; ++c) {
      double s = 0.5 * (K[r * nd + c] + K[c * nd + r]);
      K[r * nd + c] = K[c * nd + r] = s;
    }
}

// Unit-density trilinear mass matrix of an hx*hy*hz box, 2x2x2 Gauss
// (exact for products of trilinear shape functions).
static void hex_M0(const double h[3], double* M) {
  std::memset(M, 0, sizeof(double) * 64);
  std::vector<double> x, w;
  gauss_rule(2, x, w);
  const double detJ = h[0] * h[1] * h[2] / 8.0;
  for (size_t i = 0; i < x.size(); ++i)
    for (size_t j = 0; j < x.size(); ++j)
      for (size_t k = 0; k < x.size(); ++k) {
        double N[8];
        for (int a = 0; a < 8; ++a)
          N[a] = (1 + (2.0 * kCx[a] - 1) * x[i]) * (1 + (2.0 * kCy[a] - 1) * x[j]) *
                 (1 + (2.0 * kCz[a] - 1) * x[k]) / 8.0;
        const double wq = w[i] * w[j] * w[k] * detJ;
        for (int a = 0; a < 8; ++a)
          for (int b = 0; b < 8; ++b) M[a * 8 + b] += wq * N[a] * N[b];
      }
}

// Modal coefficients: Kmod = W K0 W / 64 with W[i][S] = (-1)^popcount(i & S)
// the symmetric Walsh-Hadamard matrix over corner bits (bit 0 x, bit 1 y,
// bit 2 z; W W = 8 I), modal DOF 3*S + c.
// Returns 1 and fills coef when every entry outside the box-element pattern
// vanishes (relative 1e-12), else 0 (the dense kernel is used).
static int hex_modal_coefficients(int physics, const double* K0, double* coef) {
  const int nc = physics == SKT_ELASTIC ? 3 : 1;
  const int nd = 8 * nc;
  int perm[8];  // perm[bits] = corner id
  for (int a = 0; a < 8; ++a) perm[kCx[a] | (kCy[a] << 1) | (kCz[a] << 2)] = a;
  auto h = [](int idx, int S) {  // symmetric Walsh-Hadamard sign
    return (__builtin_popcount(idx & S) & 1) ? -1.0 : 1.0;
  };
  std::vector<double> Km(nd * nd, 0.0);
  for (int S = 0; S < 8; ++S)
    for (int c = 0; c < nc; ++c)
      for (int S2 = 0; S2 < 8; ++S2)
        for (int c2 = 0; c2 < nc; ++c2) {
          double acc = 0.0;
          for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) {
              const double kk = K0[(nc * perm[i] + c) * nd + nc * perm[j] + c2];
              acc += h(i, S) * kk * h(j, S2);
            }
          Km[(nc * S + c) * nd + nc * S2 + c2] = acc / 64.0;
        }
  double mx = 0.0;
  for (double v : Km) mx = std::fabs(v) > mx ? std::fabs(v) : mx;
  std::vector<char> used(nd * nd, 0);
  auto take = [&](int i, int j) {
    used[i * nd + j] = used[j * nd + i] = 1;
    return Km[i * nd + j];
  };
  if (nc == 1) {
    // all 8 modes: mode 0 (constant) carries mass, not conduction
    for (int s = 0; s < 8; ++s) coef[s] = take(s, s);
    if (std::fabs(coef[0]) <= 1e-12 * mx) coef[0] = 0.0;
  } else {
    const int lin[3] = {3, 7, 14};
    coef[0] = take(lin[0], lin[0]); coef[1] = take(lin[1], lin[1]); coef[2] = take(lin[2], lin[2]);
    coef[3] = take(lin[0], lin[1]); coef[4] = take(lin[0], lin[2]); coef[5] = take(lin[1], lin[2]);
    const int pairs[6][2] = {{4, 6}, {5, 12}, {8, 13}, {9, 20}, {10, 17}, {19, 15}};
    for (int q = 0; q < 6; ++q) {
      const int i1 = pairs[q][0], i2 = pairs[q][1];
      coef[6 + 3 * q] = take(i1, i1);
      coef[7 + 3 * q] = take(i2, i2);
      coef[8 + 3 * q] = take(i1, i2);
    }
    const int b4[3] = {11, 18, 16};
    coef[24] = take(b4[0], b4[0]); coef[25] = take(b4[1], b4[1]); coef[26] = take(b4[2], b4[2]);
    coef[27] = take(b4[0], b4[1]); coef[28] = take(b4[0], b4[2]); coef[29] = take(b4[1], b4[2]);
    coef[30] = take(21, 21); coef[31] = take(22, 22); coef[32] = take(23, 23);
  }
  for (int i = 0; i < nd * nd; ++i)
    if (!used[i] && std::fabs(Km[i]) > 1e-12 * mx) return 0;
  return 1;
}

// ------------------------------------------------------------ hex8 kernels

// Node-centric gather: one thread per (node, column) sums the contributions
// of the (up to) 8 elements around its node.  The 27 neighbouring node values
// are loaded once into registers; K_e^0 indices are wave-uniform.
template <typename T, int NC>
__global__ void __launch_bounds__(256)
k_hex_apply(int64_t nx, int64_t ny, int64_t nz, int ncols,
            const T* __restrict__ K0, const T* __restrict__ E,
            const T* __restrict__ X, const uint8_t* __restrict__ fixed,
            T* __restrict__ Y) {
  const int64_t nnx = nx + 1, nny = ny + 1, nnz = nz + 1;
  const int64_t total = nnx * nny * nnz * ncols;
  constexpr int ND = 8 * NC;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int col = (int)(t % ncols);
    const int64_t node = t / ncols;
    const int64_t i = node % nnx, j = (node / nnx) % nny, k = node / (nnx * nny);
    T xv[27][NC];
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
      for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          const int64_t ii = i + dx - 1, jj = j + dy - 1, kk = k + dz - 1;
          const bool in = ii >= 0 && ii < nnx && jj >= 0 && jj < nny && kk >= 0 && kk < nnz;
          const int64_t nb = ii + nnx * (jj + nny * kk);
#pragma unroll
          for (int c = 0; c < NC; ++c)
            xv[dx + 3 * dy + 9 * dz][c] = in ? X[(nb * NC + c) * ncols + col] : T(0);
        }
    T acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = T(0);
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      // element slot: element (i-1+sx, j-1+sy, k-1+sz); this node is its
      // corner (1-sx, 1-sy, 1-sz)
      const int sx = kCx[s], sy = kCy[s], sz = kCz[s];
      const int64_t ex = i - 1 + sx, ey = j - 1 + sy, ez = k - 1 + sz;
      const bool ok = ex >= 0 && ex < nx && ey >= 0 && ey < ny && ez >= 0 && ez < nz;
      const T Ee = ok ? E[ex + nx * (ey + ny * ez)] : T(0);
      const int a = corner_id(1 - sx, 1 - sy, 1 - sz);
      T tmp[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) tmp[c] = T(0);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int nbi = (sx + kCx[b]) + 3 * (sy + kCy[b]) + 9 * (sz + kCz[b]);
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int d = 0; d < NC; ++d)
            tmp[c] += K0[(NC * a + c) * ND + NC * b + d] * xv[nbi][d];
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += Ee * tmp[c];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int64_t dof = node * NC + c;
      Y[dof * ncols + col] = fixed[dof] ? xv[13][c] : acc[c];
    }
  }
}

template <typename T, int NC>
__global__ void k_hex_diag(int64_t nx, int64_t ny, int64_t nz,
                           const T* __restrict__ K0, const T* __restrict__ E,
                           const uint8_t* __restrict__ fixed, T* __restrict__ diag) {
  const int64_t nnx = nx + 1, nny = ny + 1, nnz = nz + 1;
  const int64_t total = nnx * nny * nnz;
  constexpr int ND = 8 * NC;
  for (int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; node < total;
       node += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = node % nnx, j = (node / nnx) % nny, k = node / (nnx * nny);
    T acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = T(0);
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int sx = kCx[s], sy = kCy[s], sz = kCz[s];
      const int64_t ex = i - 1 + sx, ey = j - 1 + sy, ez = k - 1 + sz;
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
      const T Ee = E[ex + nx * (ey + ny * ez)];
      const int a = corner_id(1 - sx, 1 - sy, 1 - sz);
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += Ee * K0[(NC * a + c) * ND + NC * a + c];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int64_t dof = node * NC + c;
      diag[dof] = fixed[dof] ? T(1) : acc[c];
    }
  }
}

// energy[e, col] = 1/2 E_e u_e^T K0 u_e (accumulated in float64)
template <typename T, int NC>
__global__ void k_hex_energy(int64_t nx, int64_t ny, int64_t nz, int ncols,
                             const T* __restrict__ K0, const T* __restrict__ E,
                             const T* __restrict__ U, double* __restrict__ energy) {
  const int64_t nnx = nx + 1, nny = ny + 1;
  const int64_t total = nx * ny * nz * ncols;
  constexpr int ND = 8 * NC;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int col = (int)(t % ncols);
    const int64_t e = t / ncols;
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    T u[ND];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int64_t nb = (ex + kCx[b]) + nnx * ((ey + kCy[b]) + nny * (ez + kCz[b]));
#pragma unroll
      for (int d = 0; d < NC; ++d) u[NC * b + d] = U[(nb * NC + d) * ncols + col];
    }
    double q = 0.0;
#pragma unroll
    for (int r = 0; r < ND; ++r) {
      T s = T(0);
#pragma unroll
      for (int c = 0; c < ND; ++c) s += K0[r * ND + c] * u[c];
      q += (double)u[r] * (double)s;
    }
    energy[e * ncols + col] = 0.5 * (double)E[e] * q;
  }
}

// ------------------------------------------------------------ tet4 kernels

// Gradients of the barycentric coordinates (12 values, SoA) and volumes.
__global__ void k_tet_setup(int64_t n_nodes, int64_t n_el, const double* __restrict__ p,
                            const int32_t* __restrict__ conn, double* __restrict__ g,
                            double* __restrict__ vol, float* __restrict__ g32,
                            float* __restrict__ vol32, int* __restrict__ bad) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_el;
       e += (int64_t)gridDim.x * blockDim.x) {
    double x[4][3];
    for (int a = 0; a < 4; ++a) {
      const int64_t nd = conn[a * n_el + e];
      for (int d = 0; d < 3; ++d) x[a][d] = p[d * n_nodes + nd];
    }
    double J[3][3];  // columns are edges x_b - x_0
    for (int d = 0; d < 3; ++d)
      for (int b = 0; b < 3; ++b) J[d][b] = x[b + 1][d] - x[0][d];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (!(det > 0.0)) atomicAdd(bad, 1);
    const double id = 1.0 / det;
    double inv[3][3];  // rows are grad(lambda_1..3)
    inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    double gg[4][3];
    for (int d = 0; d < 3; ++d) {
      gg[1][d] = inv[0][d];
      gg[2][d] = inv[1][d];
      gg[3][d] = inv[2][d];
      gg[0][d] = -(inv[0][d] + inv[1][d] + inv[2][d]);
    }
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) {
        g[(3 * a + d) * n_el + e] = gg[a][d];
        g32[(3 * a + d) * n_el + e] = (float)gg[a][d];
      }
    vol[e] = det / 6.0;
    vol32[e] = (float)(det / 6.0);
  }
}

__global__ void k_tet_count(int64_t n_el, const int32_t* __restrict__ conn,
                            int64_t* __restrict__ cnt) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_el;
       e += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < 4; ++a)
      atomicAdd((unsigned long long*)&cnt[conn[a * n_el + e]], 1ull);
}

__global__ void k_tet_fill(int64_t n_el, const int32_t* __restrict__ conn,
                           const int64_t* __restrict__ ptr,
                           unsigned long long* __restrict__ cursor,
                           int32_t* __restrict__ nel, uint8_t* __restrict__ nloc) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_el;
       e += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < 4; ++a) {
      const int64_t nd = conn[a * n_el + e];
      const unsigned long long q = atomicAdd(&cursor[nd], 1ull);
      nel[ptr[nd] + (int64_t)q] = (int32_t)e;
      nloc[ptr[nd] + (int64_t)q] = (uint8_t)a;
    }
}

// sort each node's element list by element id (deterministic order)
__global__ void k_tet_sort(int64_t n_nodes, const int64_t* __restrict__ ptr,
                           int32_t* __restrict__ nel, uint8_t* __restrict__ nloc) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < n_nodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s0 = ptr[n], s1 = ptr[n + 1];
    for (int64_t q = s0 + 1; q < s1; ++q) {
      const int32_t v = nel[q];
      const uint8_t l = nloc[q];
      int64_t p = q - 1;
      while (p >= s0 && nel[p] > v) {
        nel[p + 1] = nel[p];
        nloc[p + 1] = nloc[p];
        --p;
      }
      nel[p + 1] = v;
      nloc[p + 1] = l;
    }
  }
}

// position of (element, corner) in the sorted node -> element lists
__global__ void k_tet_slots(int64_t n_nodes, int64_t n_el, const int64_t* __restrict__ ptr,
                            const int32_t* __restrict__ nel, const uint8_t* __restrict__ nloc,
                            int32_t* __restrict__ slot) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < n_nodes;
       n += (int64_t)gridDim.x * blockDim.x)
    for (int64_t q = ptr[n]; q < ptr[n + 1]; ++q)
      slot[(int64_t)nloc[q] * n_el + nel[q]] = (int32_t)q;
}

// Packed index of entry (a, b) of a symmetric 4 x 4 element matrix stored
// as 00 11 22 33 01 02 03 12 13 23.
__device__ __forceinline__ int sym10(int a, int b) {
  if (a == b) return a;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo == 0 ? 3 + hi : (lo == 1 ? 5 + hi : 9);
}

// Node-centric tet gather: y_a += E_e V_e B_a^T D B_e x_e over the elements
// of the node (elastic) or E_e V_e g_a . sum_b g_b x_b (conduction), plus
// the extra element matrix R_e (scalar operators with Robin terms).
template <typename T, int NC>
__global__ void __launch_bounds__(256)
k_tet_apply(int64_t n_nodes, int64_t n_el, int ncols, T lam0, T mu0, T ma, T cb,
            const int32_t* __restrict__ conn, const T* __restrict__ g,
            const T* __restrict__ vol, const int64_t* __restrict__ nptr,
            const int32_t* __restrict__ nel, const uint8_t* __restrict__ nloc,
            const T* __restrict__ E, const T* __restrict__ R, const T* __restrict__ X,
            const uint8_t* __restrict__ fixed, T* __restrict__ Y) {
  const int64_t total = n_nodes * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int col = (int)(t % ncols);
    const int64_t node = t / ncols;
    T acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = T(0);
    for (int64_t q = nptr[node]; q < nptr[node + 1]; ++q) {
      const int64_t e = nel[q];
      const int a = nloc[q];
      T gg[4][3];
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int d = 0; d < 3; ++d) gg[b][d] = g[(3 * b + d) * n_el + e];
      const T w = E[e] * vol[e];
      const T ga0 = a == 0 ? gg[0][0] : (a == 1 ? gg[1][0] : (a == 2 ? gg[2][0] : gg[3][0]));
      const T ga1 = a == 0 ? gg[0][1] : (a == 1 ? gg[1][1] : (a == 2 ? gg[2][1] : gg[3][1]));
      const T ga2 = a == 0 ? gg[0][2] : (a == 1 ? gg[1][2] : (a == 2 ? gg[2][2] : gg[3][2]));
      if (NC == 1) {
        T s = T(0), xs = T(0), xa = T(0), rx = T(0);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int64_t nb = conn[b * n_el + e];
          const T xb = X[nb * ncols + col];
          s += (ga0 * gg[b][0] + ga1 * gg[b][1] + ga2 * gg[b][2]) * xb;
          xs += xb;
          xa = b == a ? xb : xa;
          if (R) rx += R[(int64_t)sym10(a, b) * n_el + e] * xb;
        }
        // scalar operator a M + b K: P1 mass V / 20 (1 + delta_ab)
        if (ma == T(0) && cb == T(1)) acc[0] += w * s;
        else acc[0] += w * (cb * s + ma * ((xs + xa) * T(0.05)));
        if (R) acc[0] += rx;
      } else {
        T exx = 0, eyy = 0, ezz = 0, gyz = 0, gxz = 0, gxy = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int64_t nb = conn[b * n_el + e];
          const T ux = X[(nb * 3 + 0) * ncols + col];
          const T uy = X[(nb * 3 + 1) * ncols + col];
          const T uz = X[(nb * 3 + 2) * ncols + col];
          exx += gg[b][0] * ux;
          eyy += gg[b][1] * uy;
          ezz += gg[b][2] * uz;
          gyz += gg[b][2] * uy + gg[b][1] * uz;
          gxz += gg[b][2] * ux + gg[b][0] * uz;
          gxy += gg[b][1] * ux + gg[b][0] * uy;
        }
        const T tr = lam0 * (exx + eyy + ezz);
        const T sxx = tr + 2 * mu0 * exx, syy = tr + 2 * mu0 * eyy, szz = tr + 2 * mu0 * ezz;
        const T syz = mu0 * gyz, sxz = mu0 * gxz, sxy = mu0 * gxy;
        acc[0] += w * (ga0 * sxx + ga1 * sxy + ga2 * sxz);
        if (NC > 1) acc[NC > 1 ? 1 : 0] += w * (ga1 * syy + ga0 * sxy + ga2 * syz);
        if (NC > 2) acc[NC > 2 ? 2 : 0] += w * (ga2 * szz + ga0 * sxz + ga1 * syz);
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int64_t dof = node * NC + c;
      Y[dof * ncols + col] = fixed[dof] ? X[dof * ncols + col] : acc[c];
    }
  }
}

// Element-centric tet apply, phase 1: one thread per (element, column)
// reads the element's connectivity, gradients, volume and modulus once
// (structure-of-arrays, coalesced), gathers its 4 x NC values and writes its
// 4 x NC nodal forces to F[(c ncols + col) 4 n_el + slot(a, e)], each the
// same expression the node-centric kernel forms for that element.
template <typename T, int NC>
__global__ void __launch_bounds__(256)
k_tet_elem_forces(int64_t n_el, int ncols, T lam0, T mu0, T ma, T cb,
                  const int32_t* __restrict__ conn, const int32_t* __restrict__ slot,
                  const T* __restrict__ g, const T* __restrict__ vol,
                  const T* __restrict__ E, const T* __restrict__ R,
                  const T* __restrict__ X, T* __restrict__ F) {
  const int64_t ns = 4 * n_el;
  const int64_t total = n_el * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int col = (int)(t / n_el);
    const int64_t e = t - (int64_t)col * n_el;
    T gg[4][3];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int d = 0; d < 3; ++d) gg[b][d] = g[(3 * b + d) * n_el + e];
    int64_t nb[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) nb[b] = conn[b * n_el + e];
    const T w = E[e] * vol[e];
    if (NC == 1) {
      T x[4];
      T xs = T(0);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        x[b] = X[nb[b] * ncols + col];
        xs += x[b];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        T s = T(0);
#pragma unroll
        for (int b = 0; b < 4; ++b)
          s += (gg[a][0] * gg[b][0] + gg[a][1] * gg[b][1] + gg[a][2] * gg[b][2]) * x[b];
        T f;
        if (ma == T(0) && cb == T(1)) f = w * s;
        else f = w * (cb * s + ma * ((xs + x[a]) * T(0.05)));
        if (R) {
          T rx = T(0);
#pragma unroll
          for (int b = 0; b < 4; ++b) rx += R[(int64_t)sym10(a, b) * n_el + e] * x[b];
          f += rx;
        }
        F[(int64_t)col * ns + slot[a * n_el + e]] = f;
      }
    } else {
      T exx = 0, eyy = 0, ezz = 0, gyz = 0, gxz = 0, gxy = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const T ux = X[(nb[b] * 3 + 0) * ncols + col];
        const T uy = X[(nb[b] * 3 + 1) * ncols + col];
        const T uz = X[(nb[b] * 3 + 2) * ncols + col];
        exx += gg[b][0] * ux;
        eyy += gg[b][1] * uy;
        ezz += gg[b][2] * uz;
        gyz += gg[b][2] * uy + gg[b][1] * uz;
        gxz += gg[b][2] * ux + gg[b][0] * uz;
        gxy += gg[b][1] * ux + gg[b][0] * uy;
      }
      const T tr = lam0 * (exx + eyy + ezz);
      const T sxx = tr + 2 * mu0 * exx, syy = tr + 2 * mu0 * eyy, szz = tr + 2 * mu0 * ezz;
      const T syz = mu0 * gyz, sxz = mu0 * gxz, sxy = mu0 * gxy;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const T ga0 = gg[a][0], ga1 = gg[a][1], ga2 = gg[a][2];
        const int64_t q = slot[a * n_el + e];
        F[(int64_t)(0 * ncols + col) * ns + q] = w * (ga0 * sxx + ga1 * sxy + ga2 * sxz);
        F[(int64_t)(1 * ncols + col) * ns + q] = w * (ga1 * syy + ga0 * sxy + ga2 * syz);
        F[(int64_t)(2 * ncols + col) * ns + q] = w * (ga2 * szz + ga0 * sxz + ga1 * syz);
      }
    }
  }
}

// Phase 2: LPN lanes per (node, column) read the node's contiguous run of
// staged forces (lane l takes entries l, l + LPN, ...) and combine their
// partial sums by a fixed butterfly, so the result does not depend on
// timing; Dirichlet rows copy x.
template <typename T, int NC, int LPN>
__global__ void __launch_bounds__(256)
k_tet_gather(int64_t n_nodes, int64_t ns, int ncols, const int64_t* __restrict__ nptr,
             const T* __restrict__ F, const T* __restrict__ X,
             const uint8_t* __restrict__ fixed, T* __restrict__ Y) {
  static_assert(LPN >= 1 && LPN <= 64 && (LPN & (LPN - 1)) == 0, "LPN: a power of two");
  // the grid stride is a multiple of the block size, itself a multiple of
  // LPN, so the lanes of one group stay together in every pass
  const int64_t total = n_nodes * ncols * LPN;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(t & (LPN - 1));
    const int64_t grp = t / LPN;
    const int col = (int)(grp % ncols);
    const int64_t node = grp / ncols;
    T acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = T(0);
    const int64_t q1 = nptr[node + 1];
    for (int64_t q = nptr[node] + lane; q < q1; q += LPN) {
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += F[(int64_t)(c * ncols + col) * ns + q];
    }
#pragma unroll
    for (int off = LPN / 2; off >= 1; off >>= 1) {
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += __shfl_xor(acc[c], off, LPN);
    }
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int64_t dof = node * NC + c;
        Y[dof * ncols + col] = fixed[dof] ? X[dof * ncols + col] : acc[c];
      }
    }
  }
}

// Tet4 meshes on a grid of cells (six tets per cell, skt_tet4_set_grid):
// the apply without per-tet staging.  Phase 1, one thread per tet (the tet
// arrays are read with consecutive lanes on consecutive tets), computes the
// tet's four corner forces as k_tet_elem_forces does; a workgroup holds 32
// whole cells, whose six tets add their forces per cell corner in LDS (tets
// in order s, local nodes in order), and the cells' 8 corner forces are
// written corner-major, Fc[((col * 8 + corner) * NC + d) * n_cells + cell],
// so that consecutive cells are consecutive addresses.
constexpr int kTgCells = 32;

template <typename T, int NC>
__global__ void __launch_bounds__(6 * kTgCells)
k_tetg_cell_forces(int64_t n_cells, int ncols, T lam0, T mu0, T ma, T cb,
                   const int32_t* __restrict__ code, const int32_t* __restrict__ conn,
                   const T* __restrict__ g, const T* __restrict__ vol,
                   const T* __restrict__ E, const T* __restrict__ R,
                   const T* __restrict__ X, T* __restrict__ Fc) {
  constexpr int kT = 6 * kTgCells;
  __shared__ T sf[kT][4 * NC + 1];
  __shared__ int scode[kT];
  __shared__ T sacc[kTgCells][8 * NC];
  const int col = blockIdx.y;
  const int64_t n_el = 6 * n_cells;
  const int tid = threadIdx.x;
  const int64_t e = (int64_t)blockIdx.x * kT + tid;
  T f[4][NC];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < NC; ++d) f[a][d] = T(0);
  int cd = 0;
  if (e < n_el) {
    cd = code[e];
    T gg[4][3];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int d = 0; d < 3; ++d) gg[b][d] = g[(3 * b + d) * n_el + e];
    int64_t nb[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) nb[b] = conn[b * n_el + e];
    const T w = E[e] * vol[e];
    if (NC == 1) {
      T x[4];
      T xs = T(0);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        x[b] = X[nb[b] * ncols + col];
        xs += x[b];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        T s = T(0);
#pragma unroll
        for (int b = 0; b < 4; ++b)
          s += (gg[a][0] * gg[b][0] + gg[a][1] * gg[b][1] + gg[a][2] * gg[b][2]) * x[b];
        T fa;
        if (ma == T(0) && cb == T(1)) fa = w * s;
        else fa = w * (cb * s + ma * ((xs + x[a]) * T(0.05)));
        if (R) {
          T rx = T(0);
#pragma unroll
          for (int b = 0; b < 4; ++b) rx += R[(int64_t)sym10(a, b) * n_el + e] * x[b];
          fa += rx;
        }
        f[a][0] = fa;
      }
    } else {
      T exx = 0, eyy = 0, ezz = 0, gyz = 0, gxz = 0, gxy = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const T ux = X[(nb[b] * 3 + 0) * ncols + col];
        const T uy = X[(nb[b] * 3 + 1) * ncols + col];
        const T uz = X[(nb[b] * 3 + 2) * ncols + col];
        exx += gg[b][0] * ux;
        eyy += gg[b][1] * uy;
        ezz += gg[b][2] * uz;
        gyz += gg[b][2] * uy + gg[b][1] * uz;
        gxz += gg[b][2] * ux + gg[b][0] * uz;
        gxy += gg[b][1] * ux + gg[b][0] * uy;
      }
      const T tr = lam0 * (exx + eyy + ezz);
      const T sxx = tr + 2 * mu0 * exx, syy = tr + 2 * mu0 * eyy, szz = tr + 2 * mu0 * ezz;
      const T syz = mu0 * gyz, sxz = mu0 * gxz, sxy = mu0 * gxy;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const T ga0 = gg[a][0], ga1 = gg[a][1], ga2 = gg[a][2];
        f[a][0] = w * (ga0 * sxx + ga1 * sxy + ga2 * sxz);
        f[a][1] = w * (ga1 * syy + ga0 * sxy + ga2 * syz);
        f[a][2] = w * (ga2 * szz + ga0 * sxz + ga1 * syz);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < NC; ++d) sf[tid][a * NC + d] = f[a][d];
  scode[tid] = cd;
  for (int i = tid; i < kTgCells * 8 * NC; i += kT) sacc[i / (8 * NC)][i % (8 * NC)] = T(0);
  __syncthreads();
  if (tid < kTgCells) {
    for (int s = 0; s < 6; ++s) {
      const int t = 6 * tid + s;
      const int c4 = scode[t];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int corner = (c4 >> (3 * a)) & 7;
#pragma unroll
        for (int d = 0; d < NC; ++d) sacc[tid][corner * NC + d] += sf[t][a * NC + d];
      }
    }
  }
  __syncthreads();
  const int64_t cell0 = (int64_t)blockIdx.x * kTgCells;
  for (int i = tid; i < kTgCells * 8 * NC; i += kT) {
    const int cl = i % kTgCells, v = i / kTgCells;  // v = corner * NC + d
    if (cell0 + cl < n_cells)
      Fc[((int64_t)col * 8 * NC + v) * n_cells + cell0 + cl] = sacc[cl][v];
  }
}

// Phase 2: one thread per node (consecutive along x) sums the corner forces
// of its up to 8 cells, in a fixed order; Dirichlet rows copy x.
template <typename T, int NC>
__global__ void k_tetg_gather(int64_t nx, int64_t ny, int64_t nz, int ncols,
                              const T* __restrict__ Fc, const T* __restrict__ X,
                              const uint8_t* __restrict__ fixed, T* __restrict__ Y) {
  const int64_t nnx = nx + 1, nny = ny + 1, nn = nnx * nny * (nz + 1);
  const int64_t n_cells = nx * ny * nz;
  const int col = blockIdx.y;
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = n % nnx, j = (n / nnx) % nny, k = n / (nnx * nny);
    // every load of the node is unpredicated and issued before the sums (a
    // cell outside the grid reads cell 0 and is dropped by a select, which
    // keeps the skipping loop's sums): a load in a branch of its own was
    // waited for inside it
    T xv[NC];
    uint8_t fv[NC];
#pragma unroll
    for (int d = 0; d < NC; ++d) {
      const int64_t dof = n * NC + d;
      xv[d] = X[dof * ncols + col];
      fv[d] = fixed[dof];  // every tet operator holds its flags (all zero if none fixed)
    }
    T fc[8][NC];
    bool in[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int sx = q & 1, sy = (q >> 1) & 1, sz = q >> 2;  // cell (i - 1 + sx, ...)
      const int64_t ci = i - 1 + sx, cj = j - 1 + sy, ck = k - 1 + sz;
      in[q] = ci >= 0 && ci < nx && cj >= 0 && cj < ny && ck >= 0 && ck < nz;
      const int corner = (1 - sx) | ((1 - sy) << 1) | ((1 - sz) << 2);
      const int64_t cell = in[q] ? ci + nx * (cj + ny * ck) : 0;
#pragma unroll
      for (int d = 0; d < NC; ++d)
        fc[q][d] = Fc[((int64_t)col * 8 * NC + corner * NC + d) * n_cells + cell];
    }
    T acc[NC];
#pragma unroll
    for (int d = 0; d < NC; ++d) acc[d] = T(0);
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int d = 0; d < NC; ++d) {
        const T a = acc[d] + fc[q][d];
        acc[d] = in[q] ? a : acc[d];
      }
#pragma unroll
    for (int d = 0; d < NC; ++d) {
      const int64_t dof = n * NC + d;
      Y[dof * ncols + col] = fv[d] ? xv[d] : acc[d];
    }
  }
}

template <typename T, int NC>
__global__ void k_tet_diag(int64_t n_nodes, int64_t n_el, T lam0, T mu0, T ma, T cb,
                           const T* __restrict__ g, const T* __restrict__ vol,
                           const int64_t* __restrict__ nptr, const int32_t* __restrict__ nel,
                           const uint8_t* __restrict__ nloc, const T* __restrict__ E,
                           const T* __restrict__ R, const uint8_t* __restrict__ fixed,
                           T* __restrict__ diag) {
  for (int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; node < n_nodes;
       node += (int64_t)gridDim.x * blockDim.x) {
    T acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = T(0);
    for (int64_t q = nptr[node]; q < nptr[node + 1]; ++q) {
      const int64_t e = nel[q];
      const int a = nloc[q];
      const T gx = g[(3 * a + 0) * n_el + e], gy = g[(3 * a + 1) * n_el + e],
              gz = g[(3 * a + 2) * n_el + e];
      const T w = E[e] * vol[e];
      const T g2 = gx * gx + gy * gy + gz * gz;
      if (NC == 1) {
        if (ma == T(0) && cb == T(1)) acc[0] += w * g2;
        else acc[0] += w * (cb * g2 + ma * T(0.1));
        if (R) acc[0] += R[(int64_t)a * n_el + e];
      } else {
        acc[0] += w * (mu0 * g2 + (lam0 + mu0) * gx * gx);
        acc[NC > 1 ? 1 : 0] += w * (mu0 * g2 + (lam0 + mu0) * gy * gy);
        acc[NC > 2 ? 2 : 0] += w * (mu0 * g2 + (lam0 + mu0) * gz * gz);
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int64_t dof = node * NC + c;
      diag[dof] = fixed[dof] ? T(1) : acc[c];
    }
  }
}

template <typename T, int NC>
__global__ void k_tet_energy(int64_t n_el, int ncols, T lam0, T mu0,
                             const int32_t* __restrict__ conn, const T* __restrict__ g,
                             const T* __restrict__ vol, const T* __restrict__ E,
                             const T* __restrict__ U, double* __restrict__ energy) {
  const int64_t total = n_el * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int col = (int)(t % ncols);
    const int64_t e = t / ncols;
    double q;
    if (NC == 1) {
      double gx = 0, gy = 0, gz = 0;
      for (int b = 0; b < 4; ++b) {
        const double xb = U[(int64_t)conn[b * n_el + e] * ncols + col];
        gx += (double)g[(3 * b + 0) * n_el + e] * xb;
        gy += (double)g[(3 * b + 1) * n_el + e] * xb;
        gz += (double)g[(3 * b + 2) * n_el + e] * xb;
      }
      q = gx * gx + gy * gy + gz * gz;
    } else {
      double exx = 0, eyy = 0, ezz = 0, gyz = 0, gxz = 0, gxy = 0;
      for (int b = 0; b < 4; ++b) {
        const int64_t nb = conn[b * n_el + e];
        const double ux = U[(nb * 3 + 0) * ncols + col];
        const double uy = U[(nb * 3 + 1) * ncols + col];
        const double uz = U[(nb * 3 + 2) * ncols + col];
        const double bx = g[(3 * b + 0) * n_el + e], by = g[(3 * b + 1) * n_el + e],
                     bz = g[(3 * b + 2) * n_el + e];
        exx += bx * ux; eyy += by * uy; ezz += bz * uz;
        gyz += bz * uy + by * uz; gxz += bz * ux + bx * uz; gxy += by * ux + bx * uy;
      }
      const double tr = exx + eyy + ezz;
      q = (double)lam0 * tr * tr +
          2.0 * (double)mu0 * (exx * exx + eyy * eyy + ezz * ezz) +
          (double)mu0 * (gyz * gyz + gxz * gxz + gxy * gxy);
    }
    energy[e * ncols + col] = 0.5 * (double)E[e] * (double)vol[e] * q;
  }
}

// ------------------------------------------------------------ dispatch

template <typename T>
static int hex_apply_t(skt_operator* op, int ncols, const T* K0, const T* E, const T* X,
                       T* Y, hipStream_t s) {
  const int64_t total = op->n_nodes * ncols;
  if (op->ncomp == 3)
    hipLaunchKernelGGL((k_hex_apply<T, 3>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                       op->nx, op->ny, op->nz, ncols, K0, E, X, op->fixed, Y);
  else
    hipLaunchKernelGGL((k_hex_apply<T, 1>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                       op->nx, op->ny, op->nz, ncols, K0, E, X, op->fixed, Y);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

template <typename T>
static const T* tet_extra(const skt_operator* op);
template <>
const double* tet_extra<double>(const skt_operator* op) { return op->tet_x64; }
template <>
const float* tet_extra<float>(const skt_operator* op) { return op->tet_x32; }

// SKT_TET_APPLY=slots forces the per-tet staging form on grid meshes (the
// two forms sum in different orders; tests compare both with assembled
// matrices).
static bool tet_grid_apply_enabled() {
  const char* v = std::getenv("SKT_TET_APPLY");
  return !(v && std::strcmp(v, "slots") == 0);
}

template <typename T>
static int tet_apply_t(skt_operator* op, int ncols, const T* g, const T* vol, const T* E,
                       const T* X, T* Y, hipStream_t s) {
  const int64_t total = op->n_nodes * ncols;
  const T* R = tet_extra<T>(op);
  const size_t fbytes = (size_t)op->n_el * 4 * op->ncomp * ncols * sizeof(T);
  const int64_t n_cells = op->nx * op->ny * op->nz;
  const size_t cbytes = (size_t)n_cells * 8 * op->ncomp * ncols * sizeof(T);
  if (op->tet_grid && op->tet_code && op->tet_f && cbytes <= op->tet_f_bytes &&
      tet_grid_apply_enabled()) {
    // grid of cells: per-cell corner forces, then the node gather (which
    // reads the Dirichlet flags unconditionally)
    SKT_CHECK(op->fixed != nullptr, "internal: tet operator without Dirichlet flags");
    T* F = (T*)op->tet_f;
    const dim3 g1((unsigned)((n_cells + kTgCells - 1) / kTgCells), (unsigned)ncols);
    const dim3 g2(std::min<unsigned>(grid_for(op->n_nodes), 65535u), (unsigned)ncols);
    if (op->ncomp == 3) {
      hipLaunchKernelGGL((k_tetg_cell_forces<T, 3>), g1, dim3(6 * kTgCells), 0, s, n_cells,
                         ncols, (T)op->lam0, (T)op->mu0, (T)op->mass_a, (T)op->cond_b,
                         op->tet_code, op->conn, g, vol, E, R, X, F);
      hipLaunchKernelGGL((k_tetg_gather<T, 3>), g2, dim3(kBlock), 0, s, op->nx, op->ny, op->nz,
                         ncols, (const T*)F, X, op->fixed, Y);
    } else {
      hipLaunchKernelGGL((k_tetg_cell_forces<T, 1>), g1, dim3(6 * kTgCells), 0, s, n_cells,
                         ncols, (T)op->lam0, (T)op->mu0, (T)op->mass_a, (T)op->cond_b,
                         op->tet_code, op->conn, g, vol, E, R, X, F);
      hipLaunchKernelGGL((k_tetg_gather<T, 1>), g2, dim3(kBlock), 0, s, op->nx, op->ny, op->nz,
                         ncols, (const T*)F, X, op->fixed, Y);
    }
    SKT_LAUNCH_CHECK();
    return SKT_OK;
  }
  if (op->tet_f && fbytes <= op->tet_f_bytes) {
    // element forces, then the node gather (8 lanes per node and column)
    constexpr int kLpn = 8;
    T* F = (T*)op->tet_f;
    const int64_t te = op->n_el * ncols;
    const int64_t tg = total * kLpn;
    const int64_t ns = 4 * op->n_el;
    if (op->ncomp == 3) {
      hipLaunchKernelGGL((k_tet_elem_forces<T, 3>), dim3(grid_for(te)), dim3(kBlock), 0, s,
                         op->n_el, ncols, (T)op->lam0, (T)op->mu0, (T)op->mass_a,
                         (T)op->cond_b, op->conn, op->tet_slot, g, vol, E, R, X, F);
      hipLaunchKernelGGL((k_tet_gather<T, 3, kLpn>), dim3(grid_for(tg)), dim3(kBlock), 0, s,
                         op->n_nodes, ns, ncols, op->nptr, (const T*)F, X, op->fixed, Y);
    } else {
      hipLaunchKernelGGL((k_tet_elem_forces<T, 1>), dim3(grid_for(te)), dim3(kBlock), 0, s,
                         op->n_el, ncols, (T)op->lam0, (T)op->mu0, (T)op->mass_a,
                         (T)op->cond_b, op->conn, op->tet_slot, g, vol, E, R, X, F);
      hipLaunchKernelGGL((k_tet_gather<T, 1, kLpn>), dim3(grid_for(tg)), dim3(kBlock), 0, s,
                         op->n_nodes, ns, ncols, op->nptr, (const T*)F, X, op->fixed, Y);
    }
    SKT_LAUNCH_CHECK();
    return SKT_OK;
  }
  if (op->ncomp == 3)
    hipLaunchKernelGGL((k_tet_apply<T, 3>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                       op->n_nodes, op->n_el, ncols, (T)op->lam0, (T)op->mu0, (T)op->mass_a,
                       (T)op->cond_b, op->conn, g,
                       vol, op->nptr, op->nel, op->nloc, E, R, X, op->fixed, Y);
  else
    hipLaunchKernelGGL((k_tet_apply<T, 1>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                       op->n_nodes, op->n_el, ncols, (T)op->lam0, (T)op->mu0, (T)op->mass_a,
                       (T)op->cond_b, op->conn, g,
                       vol, op->nptr, op->nel, op->nloc, E, R, X, op->fixed, Y);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int op_apply(skt_operator* op, int dtype, int ncols, int kc, const void* E, const void* X,
             void* Y, hipStream_t s) {
  SKT_CHECK(op && E && X && Y, "null pointer");
  SKT_CHECK(ncols >= 1 && ncols <= 64, "ncols must be in [1, 64]");
  SKT_CHECK(kc >= 1 && ncols % kc == 0, "column block must divide ncols");
  SKT_CHECK(X != Y, "operator apply cannot run in place");
  SKT_CHECK(dtype == SKT_F64 || dtype == SKT_F32, "dtype must be SKT_F64 or SKT_F32");
  if (op->mesh == SKT_MESH_HEX8 && op->modal_ok) {
    int st;
    if (dtype == SKT_F64)
      st = hex_apply_modal<double>(op, ncols, kc, op->modal_d, (const double*)E,
                                   (const double*)X, (double*)Y, s);
    else
      st = hex_apply_modal<float>(op, ncols, kc, op->modal_f, (const float*)E,
                                  (const float*)X, (float*)Y, s);
    if (st) return st;
    return extra_apply_add(op, dtype, ncols, kc, X, Y, s);
  }
  SKT_CHECK(kc == ncols, "column-blocked vectors need the structured hex8 operator");
  if (op->mesh == SKT_MESH_HEX8) {
    int st;
    if (dtype == SKT_F64)
      st = hex_apply_t<double>(op, ncols, op->K0d, (const double*)E, (const double*)X,
                               (double*)Y, s);
    else
      st = hex_apply_t<float>(op, ncols, op->K0f, (const float*)E, (const float*)X,
                              (float*)Y, s);
    if (st) return st;
    return extra_apply_add(op, dtype, ncols, kc, X, Y, s);
  }
  if (dtype == SKT_F64)
    return tet_apply_t<double>(op, ncols, op->g64, op->vol64, (const double*)E,
                               (const double*)X, (double*)Y, s);
  return tet_apply_t<float>(op, ncols, op->g32, op->vol32, (const float*)E,
                            (const float*)X, (float*)Y, s);
}

int op_apply_dot(skt_operator* op, int dtype, int ncols, int kc, const void* E, const void* X,
                 void* Y, double* dot, int max_slots, int* slots, hipStream_t s,
                 const double* stop) {
  SKT_CHECK(slots, "null pointer");
  *slots = 0;
  const bool fuse = op && dot && op->mesh == SKT_MESH_HEX8 && op->modal_ok && !op->extra64 &&
                    ncols >= 1 && kc >= 1 && ncols % kc == 0 &&
                    hex_modal_dot_slots(op, ncols, kc) <= max_slots;
  if (!fuse) return op_apply(op, dtype, ncols, kc, E, X, Y, s);
  SKT_CHECK(E && X && Y, "null pointer");
  SKT_CHECK(ncols <= 64, "ncols must be in [1, 64]");
  SKT_CHECK(X != Y, "operator apply cannot run in place");
  SKT_CHECK(dtype == SKT_F64 || dtype == SKT_F32, "dtype must be SKT_F64 or SKT_F32");
  const int st = dtype == SKT_F64
                     ? hex_apply_modal<double>(op, ncols, kc, op->modal_d, (const double*)E,
                                               (const double*)X, (double*)Y, s, dot, stop)
                     : hex_apply_modal<float>(op, ncols, kc, op->modal_f, (const float*)E,
                                              (const float*)X, (float*)Y, s, dot, stop);
  if (st == SKT_OK) *slots = hex_modal_dot_slots(op, ncols, kc);
  return st;
}

template <typename T>
static int diag_t(skt_operator* op, const T* K0, const T* g, const T* vol, const T* E,
                  T* d, hipStream_t s) {
  if (op->mesh == SKT_MESH_HEX8) {
    if (op->ncomp == 3)
      hipLaunchKernelGGL((k_hex_diag<T, 3>), dim3(grid_for(op->n_nodes)), dim3(kBlock), 0, s,
                         op->nx, op->ny, op->nz, K0, E, op->fixed, d);
    else
      hipLaunchKernelGGL((k_hex_diag<T, 1>), dim3(grid_for(op->n_nodes)), dim3(kBlock), 0, s,
                         op->nx, op->ny, op->nz, K0, E, op->fixed, d);
  } else {
    if (op->ncomp == 3)
      hipLaunchKernelGGL((k_tet_diag<T, 3>), dim3(grid_for(op->n_nodes)), dim3(kBlock), 0, s,
                         op->n_nodes, op->n_el, (T)op->lam0, (T)op->mu0, (T)op->mass_a,
                         (T)op->cond_b, g, vol, op->nptr,
                         op->nel, op->nloc, E, tet_extra<T>(op), op->fixed, d);
    else
      hipLaunchKernelGGL((k_tet_diag<T, 1>), dim3(grid_for(op->n_nodes)), dim3(kBlock), 0, s,
                         op->n_nodes, op->n_el, (T)op->lam0, (T)op->mu0, (T)op->mass_a,
                         (T)op->cond_b, g, vol, op->nptr,
                         op->nel, op->nloc, E, tet_extra<T>(op), op->fixed, d);
  }
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int op_diag(skt_operator* op, int dtype, const void* E, void* diag, hipStream_t s) {
  SKT_CHECK(op && E && diag, "null pointer");
  int st;
  if (dtype == SKT_F64) {
    st = diag_t<double>(op, op->K0d, op->g64, op->vol64, (const double*)E, (double*)diag, s);
  } else {
    SKT_CHECK(dtype == SKT_F32, "dtype must be SKT_F64 or SKT_F32");
    st = diag_t<float>(op, op->K0f, op->g32, op->vol32, (const float*)E, (float*)diag, s);
  }
  if (st) return st;
  return extra_diag_add(op, dtype, diag, s);
}

template <typename T>
static int energy_t(skt_operator* op, int ncols, const T* K0, const T* g, const T* vol,
                    const T* E, const T* U, double* out, hipStream_t s) {
  const int64_t total = op->n_el * ncols;
  if (op->mesh == SKT_MESH_HEX8 && op->modal_ok) {
    const T* coef;
    if constexpr (sizeof(T) == 8) coef = op->modal_d;
    else coef = op->modal_f;
    return hex_energy_modal<T>(op, ncols, coef, E, U, out, s);
  }
  if (op->mesh == SKT_MESH_HEX8) {
    if (op->ncomp == 3)
      hipLaunchKernelGGL((k_hex_energy<T, 3>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                         op->nx, op->ny, op->nz, ncols, K0, E, U, out);
    else
      hipLaunchKernelGGL((k_hex_energy<T, 1>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                         op->nx, op->ny, op->nz, ncols, K0, E, U, out);
  } else {
    if (op->ncomp == 3)
      hipLaunchKernelGGL((k_tet_energy<T, 3>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                         op->n_el, ncols, (T)op->lam0, (T)op->mu0, op->conn, g, vol, E, U, out);
    else
      hipLaunchKernelGGL((k_tet_energy<T, 1>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                         op->n_el, ncols, (T)op->lam0, (T)op->mu0, op->conn, g, vol, E, U, out);
  }
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

}  // namespace skt

using namespace skt;

extern "C" {

int skt_hex8_create(int physics, int64_t nx, int64_t ny, int64_t nz, double hx,
                    double hy, double hz, double nu, int quad_order,
                    const uint8_t* dof_fixed, skt_operator** out) {
  SKT_CHECK(out, "null output handle");
  *out = nullptr;
  SKT_CHECK(physics == SKT_ELASTIC || physics == SKT_HEAT, "physics must be elastic or heat");
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(hx > 0 && hy > 0 && hz > 0, "element sizes must be positive");
  SKT_CHECK(quad_order >= 1 && quad_order <= 4, "quad_order must be in [1, 4]");
  SKT_CHECK(physics == SKT_HEAT || (nu > -1.0 && nu < 0.5), "nu must be in (-1, 0.5)");
  SKT_CHECK(dof_fixed, "dof_fixed is required");
  SKT_CHECK((nx + 1) * (ny + 1) * (nz + 1) * 3 < (int64_t)INT32_MAX, "grid too large");
  skt_operator* op = new skt_operator();
  op->mesh = SKT_MESH_HEX8;
  op->physics = physics;
  op->ncomp = physics == SKT_ELASTIC ? 3 : 1;
  op->nx = nx; op->ny = ny; op->nz = nz;
  op->h[0] = hx; op->h[1] = hy; op->h[2] = hz;
  op->n_nodes = (nx + 1) * (ny + 1) * (nz + 1);
  op->n_el = nx * ny * nz;
  op->ndof = op->n_nodes * op->ncomp;
  hex_K0(physics, nu, op->h, quad_order, op->K0h);
  return hex8_finish(op, dof_fixed, out);
}

int skt_hex8_create_scalar(int64_t nx, int64_t ny, int64_t nz, double hx, double hy, double hz,
                           double mass_coef, double stiff_coef, const uint8_t* dof_fixed,
                           skt_operator** out) {
  SKT_CHECK(out, "null output handle");
  *out = nullptr;
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(hx > 0 && hy > 0 && hz > 0, "element sizes must be positive");
  SKT_CHECK(mass_coef >= 0 && stiff_coef >= 0 && mass_coef + stiff_coef > 0,
            "coefficients must be >= 0 and not both zero");
  SKT_CHECK(dof_fixed, "dof_fixed is required");
  SKT_CHECK((nx + 1) * (ny + 1) * (nz + 1) < (int64_t)INT32_MAX, "grid too large");
  skt_operator* op = new skt_operator();
  op->mesh = SKT_MESH_HEX8;
  op->physics = SKT_HEAT;
  op->ncomp = 1;
  op->nx = nx; op->ny = ny; op->nz = nz;
  op->h[0] = hx; op->h[1] = hy; op->h[2] = hz;
  op->n_nodes = (nx + 1) * (ny + 1) * (nz + 1);
  op->n_el = nx * ny * nz;
  op->ndof = op->n_nodes;
  double Kc[64], M[64];
  hex_K0(SKT_HEAT, 0.0, op->h, 2, Kc);
  hex_M0(op->h, M);
  for (int i = 0; i < 64; ++i) op->K0h[i] = mass_coef * M[i] + stiff_coef * Kc[i];
  return hex8_finish(op, dof_fixed, out);
}

}  // extern "C"

namespace skt {
// Shared tail of the hex8 constructors: modal coefficients, device copies
// of K_e^0, Dirichlet mask and list.
int hex8_finish(skt_operator* op, const uint8_t* dof_fixed, skt_operator** out) {
  const int physics = op->physics;
  const int nd = 8 * op->ncomp;
  op->modal_ok = hex_modal_coefficients(physics, op->K0h, op->modal_d);
  for (int i = 0; i < 40; ++i) op->modal_f[i] = (float)op->modal_d[i];
  std::vector<float> kf(nd * nd);
  for (int i = 0; i < nd * nd; ++i) kf[i] = (float)op->K0h[i];
  SKT_HIP(hipMalloc(&op->K0d, sizeof(double) * nd * nd));
  SKT_HIP(hipMalloc(&op->K0f, sizeof(float) * nd * nd));
  SKT_HIP(hipMemcpy(op->K0d, op->K0h, sizeof(double) * nd * nd, hipMemcpyHostToDevice));
  SKT_HIP(hipMemcpy(op->K0f, kf.data(), sizeof(float) * nd * nd, hipMemcpyHostToDevice));
  SKT_HIP(hipMalloc(&op->fixed, op->ndof));
  SKT_HIP(hipMemcpy(op->fixed, dof_fixed, op->ndof, hipMemcpyDeviceToDevice));
  {
    std::vector<uint8_t> hf(op->ndof);
    SKT_HIP(hipMemcpy(hf.data(), dof_fixed, op->ndof, hipMemcpyDeviceToHost));
    std::vector<int32_t> lst;
    for (int64_t i = 0; i < op->ndof; ++i)
      if (hf[i]) lst.push_back((int32_t)i);
    op->n_fixed = (int64_t)lst.size();
    if (!lst.empty()) {
      SKT_HIP(hipMalloc(&op->fixed_list, sizeof(int32_t) * lst.size()));
      SKT_HIP(hipMemcpy(op->fixed_list, lst.data(), sizeof(int32_t) * lst.size(),
                        hipMemcpyHostToDevice));
    }
  }
  *out = op;
  return SKT_OK;
}
}  // namespace skt

extern "C" {

int skt_tet4_create(int physics, int64_t n_nodes, int64_t n_el, const double* coords,
                    const int32_t* conn, double nu, const uint8_t* dof_fixed,
                    skt_operator** out) {
  SKT_CHECK(out, "null output handle");
  *out = nullptr;
  SKT_CHECK(physics == SKT_ELASTIC || physics == SKT_HEAT, "physics must be elastic or heat");
  SKT_CHECK(n_nodes >= 4 && n_el >= 1 && n_el < (int64_t)INT32_MAX, "bad mesh size");
  SKT_CHECK(coords && conn && dof_fixed, "null pointer");
  SKT_CHECK(physics == SKT_HEAT || (nu > -1.0 && nu < 0.5), "nu must be in (-1, 0.5)");
  skt_operator* op = new skt_operator();
  op->mesh = SKT_MESH_TET4;
  op->physics = physics;
  op->ncomp = physics == SKT_ELASTIC ? 3 : 1;
  op->n_nodes = n_nodes;
  op->n_el = n_el;
  op->ndof = n_nodes * op->ncomp;
  op->lam0 = physics == SKT_ELASTIC ? nu / ((1.0 + nu) * (1.0 - 2.0 * nu)) : 0.0;
  op->mu0 = physics == SKT_ELASTIC ? 1.0 / (2.0 * (1.0 + nu)) : 0.0;
  // connectivity range check on the host (setup only)
  std::vector<int32_t> hc(4 * n_el);
  SKT_HIP(hipMemcpy(hc.data(), conn, sizeof(int32_t) * 4 * n_el, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < 4 * n_el; ++i)
    SKT_CHECK(hc[i] >= 0 && hc[i] < n_nodes, "connectivity index out of range");
  SKT_HIP(hipMalloc(&op->conn, sizeof(int32_t) * 4 * n_el));
  SKT_HIP(hipMemcpy(op->conn, conn, sizeof(int32_t) * 4 * n_el, hipMemcpyDeviceToDevice));
  SKT_HIP(hipMalloc(&op->g64, sizeof(double) * 12 * n_el));
  SKT_HIP(hipMalloc(&op->g32, sizeof(float) * 12 * n_el));
  SKT_HIP(hipMalloc(&op->vol64, sizeof(double) * n_el));
  SKT_HIP(hipMalloc(&op->vol32, sizeof(float) * n_el));
  int* bad = nullptr;
  SKT_HIP(hipMalloc(&bad, sizeof(int)));
  SKT_HIP(hipMemset(bad, 0, sizeof(int)));
  hipLaunchKernelGGL(k_tet_setup, dim3(grid_for(n_el)), dim3(kBlock), 0, 0, n_nodes, n_el,
                     coords, op->conn, op->g64, op->vol64, op->g32, op->vol32, bad);
  SKT_LAUNCH_CHECK();
  int hbad = 0;
  SKT_HIP(hipMemcpy(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost));
  (void)hipFree(bad);
  SKT_CHECK(hbad == 0, "tetrahedron with non-positive volume");
  // node -> element lists
  int64_t* cnt = nullptr;
  SKT_HIP(hipMalloc(&cnt, sizeof(int64_t) * (n_nodes + 1)));
  SKT_HIP(hipMemset(cnt, 0, sizeof(int64_t) * (n_nodes + 1)));
  SKT_HIP(hipMalloc(&op->nptr, sizeof(int64_t) * (n_nodes + 1)));
  hipLaunchKernelGGL(k_tet_count, dim3(grid_for(n_el)), dim3(kBlock), 0, 0, n_el, op->conn, cnt);
  SKT_LAUNCH_CHECK();
  size_t bytes = 0;
  SKT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt, op->nptr, (int)(n_nodes + 1)));
  void* tmp = nullptr;
  SKT_HIP(hipMalloc(&tmp, bytes ? bytes : 16));
  SKT_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, op->nptr, (int)(n_nodes + 1)));
  SKT_HIP(hipDeviceSynchronize());
  (void)hipFree(tmp);
  SKT_HIP(hipMemset(cnt, 0, sizeof(int64_t) * (n_nodes + 1)));
  SKT_HIP(hipMalloc(&op->nel, sizeof(int32_t) * 4 * n_el));
  SKT_HIP(hipMalloc(&op->nloc, 4 * n_el));
  hipLaunchKernelGGL(k_tet_fill, dim3(grid_for(n_el)), dim3(kBlock), 0, 0, n_el, op->conn,
                     op->nptr, (unsigned long long*)cnt, op->nel, op->nloc);
  hipLaunchKernelGGL(k_tet_sort, dim3(grid_for(n_nodes)), dim3(kBlock), 0, 0, n_nodes,
                     op->nptr, op->nel, op->nloc);
  SKT_LAUNCH_CHECK();
  SKT_HIP(hipDeviceSynchronize());
  (void)hipFree(cnt);
  SKT_HIP(hipMalloc(&op->fixed, op->ndof));
  SKT_HIP(hipMemcpy(op->fixed, dof_fixed, op->ndof, hipMemcpyDeviceToDevice));
  {
    std::vector<uint8_t> hf(op->ndof);
    SKT_HIP(hipMemcpy(hf.data(), dof_fixed, op->ndof, hipMemcpyDeviceToHost));
    int64_t nf = 0;
    for (int64_t i = 0; i < op->ndof; ++i) nf += hf[i] ? 1 : 0;
    op->n_fixed = nf;
  }
  // element-force staging of the apply: one float64 column (two float32),
  // in slot order (slots are int32: meshes with 4 n_el >= 2^31 use the
  // node-centric apply)
  if (4 * n_el < (int64_t)INT32_MAX) {
    SKT_HIP(hipMalloc(&op->tet_slot, sizeof(int32_t) * 4 * n_el));
    hipLaunchKernelGGL(k_tet_slots, dim3(grid_for(n_nodes)), dim3(kBlock), 0, 0, n_nodes, n_el,
                       op->nptr, op->nel, op->nloc, op->tet_slot);
    SKT_LAUNCH_CHECK();
    SKT_HIP(hipDeviceSynchronize());
    op->tet_f_bytes = (size_t)n_el * 4 * op->ncomp * sizeof(double);
    SKT_HIP(hipMalloc(&op->tet_f, op->tet_f_bytes));
  }
  *out = op;
  return SKT_OK;
}

int skt_tet4_set_grid(skt_operator* op, int64_t nx, int64_t ny, int64_t nz) {
  SKT_CHECK(op && op->mesh == SKT_MESH_TET4, "a tet4 operator is required");
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  const int64_t nnx = nx + 1, nny = ny + 1;
  SKT_CHECK(op->n_nodes == nnx * nny * (nz + 1) && op->n_el == 6 * nx * ny * nz,
            "the mesh is not a grid of nx*ny*nz cells split into six tets each");
  // every tet 6 c + s has its four nodes among the corners of cell c
  std::vector<int32_t> hc(4 * op->n_el);
  SKT_HIP(hipMemcpy(hc.data(), op->conn, sizeof(int32_t) * 4 * op->n_el,
                    hipMemcpyDeviceToHost));
  for (int64_t e = 0; e < op->n_el; ++e) {
    const int64_t c = e / 6;
    const int64_t cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    for (int q = 0; q < 4; ++q) {
      const int64_t n = hc[(size_t)q * op->n_el + e];
      const int64_t i = n % nnx, j = (n / nnx) % nny, k = n / (nnx * nny);
      SKT_CHECK(i - cx >= 0 && i - cx <= 1 && j - cy >= 0 && j - cy <= 1 && k - cz >= 0 &&
                    k - cz <= 1,
                "a tet's nodes are not corners of its grid cell (element 6 c + s in cell c)");
    }
  }
  // each tet's local nodes as corners of its cell
  std::vector<int32_t> code(op->n_el);
  for (int64_t e = 0; e < op->n_el; ++e) {
    const int64_t c = e / 6;
    const int64_t cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    int32_t v = 0;
    for (int q = 0; q < 4; ++q) {
      const int64_t n = hc[(size_t)q * op->n_el + e];
      const int64_t i = n % nnx, j = (n / nnx) % nny, k = n / (nnx * nny);
      v |= (int32_t)((i - cx) | ((j - cy) << 1) | ((k - cz) << 2)) << (3 * q);
    }
    code[(size_t)e] = v;
  }
  if (!op->tet_code) SKT_HIP(hipMalloc(&op->tet_code, sizeof(int32_t) * op->n_el));
  SKT_HIP(hipMemcpy(op->tet_code, code.data(), sizeof(int32_t) * op->n_el,
                    hipMemcpyHostToDevice));
  op->nx = nx;
  op->ny = ny;
  op->nz = nz;
  op->tet_grid = 1;
  return SKT_OK;
}

int skt_tet4_create_scalar(int64_t n_nodes, int64_t n_el, const double* coords,
                           const int32_t* conn, double a_mass, double b_cond,
                           const uint8_t* node_fixed, skt_operator** out) {
  SKT_CHECK(out, "null output handle");
  SKT_CHECK(a_mass >= 0.0 && b_cond >= 0.0 && a_mass + b_cond > 0.0,
            "a_mass and b_cond must be >= 0 and not both 0");
  int st = skt_tet4_create(SKT_HEAT, n_nodes, n_el, coords, conn, 0.0, node_fixed, out);
  if (st) return st;
  (*out)->mass_a = a_mass;
  (*out)->cond_b = b_cond;
  return SKT_OK;
}

int skt_operator_destroy(skt_operator* op) {
  if (!op) return SKT_OK;
  (void)hipFree(op->K0d);
  (void)hipFree(op->K0f);
  (void)hipFree(op->conn);
  (void)hipFree(op->g64);
  (void)hipFree(op->g32);
  (void)hipFree(op->vol64);
  (void)hipFree(op->vol32);
  (void)hipFree(op->nptr);
  (void)hipFree(op->nel);
  (void)hipFree(op->nloc);
  (void)hipFree(op->fixed);
  (void)hipFree(op->fixed_list);
  (void)hipFree(op->extra64);
  (void)hipFree(op->extra32);
  (void)hipFree(op->tet_f);
  (void)hipFree(op->tet_slot);
  (void)hipFree(op->tet_code);
  (void)hipFree(op->tet_x64);
  (void)hipFree(op->tet_x32);
  delete op;
  return SKT_OK;
}

int skt_operator_info(const skt_operator* op, int64_t* ndof, int64_t* n_el, int* ncomp) {
  SKT_CHECK(op, "null operator");
  if (ndof) *ndof = op->ndof;
  if (n_el) *n_el = op->n_el;
  if (ncomp) *ncomp = op->ncomp;
  return SKT_OK;
}

int skt_operator_element_matrix(const skt_operator* op, double* host_out) {
  SKT_CHECK(op && host_out, "null pointer");
  SKT_CHECK(op->mesh == SKT_MESH_HEX8, "element matrix is only shared on hex8 grids");
  const int nd = 8 * op->ncomp;
  std::memcpy(host_out, op->K0h, sizeof(double) * nd * nd);
  return SKT_OK;
}

int skt_operator_apply(skt_operator* op, int dtype, int ncols, const void* E,
                       const void* X, void* Y, hipStream_t stream) {
  return op_apply(op, dtype, ncols, ncols, E, X, Y, stream);
}

int skt_operator_diag(skt_operator* op, int dtype, const void* E, void* diag,
                      hipStream_t stream) {
  return op_diag(op, dtype, E, diag, stream);
}

int skt_operator_energy(skt_operator* op, int dtype, int ncols, const void* E,
                        const void* U, double* energy, hipStream_t stream) {
  SKT_CHECK(op && E && U && energy, "null pointer");
  SKT_CHECK(ncols >= 1 && ncols <= 64, "ncols must be in [1, 64]");
  if (dtype == SKT_F64)
    return energy_t<double>(op, ncols, op->K0d, op->g64, op->vol64, (const double*)E,
                            (const double*)U, energy, stream);
  SKT_CHECK(dtype == SKT_F32, "dtype must be SKT_F64 or SKT_F32");
  return energy_t<float>(op, ncols, op->K0f, op->g32, op->vol32, (const float*)E,
                         (const float*)U, energy, stream);
}

}  // extern "C"
