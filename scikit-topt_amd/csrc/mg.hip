// Geometric multigrid preconditioner for structured hex8 operators: level
// planning, Galerkin coarsening, stencil assembly, coarsest dense inverse,
// smoother bounds and the V-cycle.  Every reduction runs in a fixed order,
// so a V-cycle is bit-reproducible.
//
// Element-matrix indexing inside this file: corner bits (bit 0 x, bit 1 y,
// bit 2 z), local DOF = NC * bits + component.  Vectors use the solver's
// column-blocked layout: value (dof, c) at ((c / kc) * ndof + dof) * kc + c % kc.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>

#include "mg.hpp"

namespace skt {

// skfem corner order of the operator's K_e^0 (see operators.hip)
static int corner_id_h(int cx, int cy, int cz) {
  return cz * 4 + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0));
}

static int axis_factor(int64_t n) {
  for (int f : {2, 3, 5})
    if (n % f == 0 && n / f >= 1) return f;
  return 1;
}

bool mg_plan(const skt_operator* op, const MGConfig& cfg, std::vector<MGLevel>* out) {
  out->clear();
  if (op->mesh != SKT_MESH_HEX8 && !(op->mesh == SKT_MESH_TET4 && op->tet_grid)) return false;
  const int NC = op->ncomp;
  MGLevel l0;
  l0.nx = op->nx; l0.ny = op->ny; l0.nz = op->nz;
  l0.nn = (l0.nx + 1) * (l0.ny + 1) * (l0.nz + 1);
  l0.ndof = l0.nn * NC;
  out->push_back(l0);
  while (out->back().ndof > cfg.max_coarse) {
    const MGLevel& f = out->back();
    int fac[3] = {axis_factor(f.nx), axis_factor(f.ny), axis_factor(f.nz)};
    // an axis is not collapsed to a single element while a longer axis
    // (more than 5 elements) still coarsens: e.g. 15x5x5 -> 5x5x5, not 5x1x1
    const int64_t n3[3] = {f.nx, f.ny, f.nz};
    bool longer = false;
    for (int a = 0; a < 3; ++a) longer = longer || (n3[a] > 5 && fac[a] > 1);
    if (longer)
      for (int a = 0; a < 3; ++a)
        if (fac[a] > 1 && n3[a] / fac[a] == 1) fac[a] = 1;
    if (fac[0] == 1 && fac[1] == 1 && fac[2] == 1) return false;
    MGLevel c;
    c.nx = f.nx / fac[0]; c.ny = f.ny / fac[1]; c.nz = f.nz / fac[2];
    c.f[0] = fac[0]; c.f[1] = fac[1]; c.f[2] = fac[2];
    c.nn = (c.nx + 1) * (c.ny + 1) * (c.nz + 1);
    c.ndof = c.nn * NC;
    out->push_back(c);
  }
  return out->size() >= 2;
}

// ------------------------------------------------------------- Galerkin

// Trilinear weight of coarse corner Q (bits) at fine corner q of child
// (a, b, c) inside a coarse element divided by (fx, fy, fz).
__device__ __forceinline__ double tri_weight(int q, int Q, int a, int b, int c, int fx, int fy,
                                             int fz) {
  const double sx = (double)(a + (q & 1)) / fx;
  const double sy = (double)(b + ((q >> 1) & 1)) / fy;
  const double sz = (double)(c + ((q >> 2) & 1)) / fz;
  const double wx = (Q & 1) ? sx : 1.0 - sx;
  const double wy = (Q & 2) ? sy : 1.0 - sy;
  const double wz = (Q & 4) ? sz : 1.0 - sz;
  return wx * wy * wz;
}

// acc += P_j^T sK P_j for child (a, b, c) of a coarse element divided by
// (fx, fy, fz); sK holds the child's ND x ND matrix, 64 threads, acc[k] is
// entry tid + 64 k.
template <int NC>
__device__ __forceinline__ void galerkin_child(const double* sK, double* sT, double* sP, int a,
                                               int b, int c, int fx, int fy, int fz,
                                               double* acc) {
  constexpr int ND = 8 * NC;
  constexpr int NN = ND * ND;
  constexpr int PER = (NN + 63) / 64;
  const int tid = threadIdx.x;
  sP[tid] = tri_weight(tid >> 3, tid & 7, a, b, c, fx, fy, fz);  // sP[q*8+Q]
  __syncthreads();
  // T[(q,d)][(R,e)] = sum_r K[(q,d)][(r,e)] P[r][R]
  for (int i = tid; i < NN; i += 64) {
    const int row = i / ND, col = i % ND;
    const int R = col / NC, e = col % NC;
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) t += sK[row * ND + r * NC + e] * sP[r * 8 + R];
    sT[i] = t;
  }
  __syncthreads();
  // Kc[(Q,d)][(R,e)] += sum_q P[q][Q] T[(q,d)][(R,e)]
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = tid + 64 * k;
    if (i < NN) {
      const int row = i / ND, col = i % ND;
      const int Q = row / NC, d = row % NC;
      double t = 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) t += sP[q * 8 + Q] * sT[(q * NC + d) * ND + col];
      acc[k] += t;
    }
  }
  __syncthreads();
}

// One workgroup of 64 threads per coarse element:
//   Kc = sum_children P_j^T K_j P_j,
// K_j = E_j Z_j K0 Z_j on the finest level (Z_j removes Dirichlet DOFs), or
// the finer level's Galerkin element matrix.  With `only`, elements whose
// flag is 0 are left to k_galerkin_interior.
template <int NC>
__global__ void __launch_bounds__(64)
k_galerkin(int64_t ncx, int64_t ncy, int64_t ncz, int fx, int fy, int fz,
           const double* __restrict__ Kf, const double* __restrict__ K0b,
           const double* __restrict__ E, const uint8_t* __restrict__ fixed,
           const uint8_t* __restrict__ only, double* __restrict__ Kc) {
  constexpr int ND = 8 * NC;
  constexpr int NN = ND * ND;
  constexpr int PER = (NN + 63) / 64;
  __shared__ double sK[NN];
  __shared__ double sT[NN];
  __shared__ double sP[64];
  const int64_t ce = blockIdx.x;
  if (only && !only[ce]) return;
  const int64_t CX = ce % ncx, CY = (ce / ncx) % ncy, CZ = ce / (ncx * ncy);
  const int64_t nfx = ncx * fx, nfy = ncy * fy;
  const int64_t nfnx = nfx + 1, nfny = nfy + 1;
  const int tid = threadIdx.x;
  double acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = 0.0;
  for (int c = 0; c < fz; ++c)
    for (int b = 0; b < fy; ++b)
      for (int a = 0; a < fx; ++a) {
        const int64_t ex = CX * fx + a, ey = CY * fy + b, ez = CZ * fz + c;
        const int64_t fe = ex + nfx * (ey + nfy * ez);
        // child matrix
        if (Kf) {
          for (int i = tid; i < NN; i += 64) sK[i] = Kf[fe * NN + i];
        } else {
          const double Ee = E[fe];
          for (int i = tid; i < NN; i += 64) {
            const int r = i / ND, s = i % ND;
            double v = Ee * K0b[i];
            if (fixed) {
              const int qr = r / NC, qs = s / NC;
              const int64_t nr = (ex + (qr & 1)) + nfnx * ((ey + ((qr >> 1) & 1)) +
                                                           nfny * (ez + (qr >> 2)));
              const int64_t ns = (ex + (qs & 1)) + nfnx * ((ey + ((qs >> 1) & 1)) +
                                                           nfny * (ez + (qs >> 2)));
              if (fixed[nr * NC + r % NC] || fixed[ns * NC + s % NC]) v = 0.0;
            }
            sK[i] = v;
          }
        }
        galerkin_child<NC>(sK, sT, sP, a, b, c, fx, fy, fz, acc);
      }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = tid + 64 * k;
    if (i < NN) Kc[ce * NN + i] = acc[k];
  }
}

// First coarse level of a tet4 operator whose mesh is a grid of nx*ny*nz
// cells, each split into six tets (element 6 c + s lies in cell c, nodes
// numbered as the cell corners of the grid).  One workgroup of 64 threads
// per coarse element: every child cell's 8-corner matrix
//   K_cell = sum_{s < 6} E_t Z_t K_t Z_t   (t = 6 cell + s, in order s)
// is gathered in LDS (each tet's K_t = V B^T D B from its barycentric
// gradients, Z removing Dirichlet DOFs) and projected as the hex children.
template <int NC>
__global__ void __launch_bounds__(64)
k_galerkin_tet(int64_t ncx, int64_t ncy, int64_t ncz, int fx, int fy, int fz,
               const int32_t* __restrict__ conn, const double* __restrict__ g,
               const double* __restrict__ vol, double lam0, double mu0, int64_t n_el,
               const double* __restrict__ E, const uint8_t* __restrict__ fixed,
               double* __restrict__ Kc) {
  constexpr int ND = 8 * NC;
  constexpr int NN = ND * ND;
  constexpr int PER = (NN + 63) / 64;
  __shared__ double sK[NN];
  __shared__ double sT[NN];
  __shared__ double sP[64];
  __shared__ double sg[6][12];    // barycentric gradients of the cell's six tets
  __shared__ int scor[6][4];      // per tet: local node -> cell corner bits, or -1
  __shared__ double sw[6];        // per tet: E_t V_t
  __shared__ int64_t snode[8];    // cell corner bits -> global node
  __shared__ uint8_t sfix[ND];    // the cell corners' Dirichlet flags
  const int64_t ce = blockIdx.x;
  const int64_t CX = ce % ncx, CY = (ce / ncx) % ncy, CZ = ce / (ncx * ncy);
  const int64_t nfx = ncx * fx, nfy = ncy * fy;
  const int64_t nfnx = nfx + 1, nfny = nfy + 1;
  const int tid = threadIdx.x;
  double acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = 0.0;
  for (int c = 0; c < fz; ++c)
    for (int b = 0; b < fy; ++b)
      for (int a = 0; a < fx; ++a) {
        const int64_t ex = CX * fx + a, ey = CY * fy + b, ez = CZ * fz + c;
        const int64_t cell = ex + nfx * (ey + nfy * ez);
        // (the previous child's projection ended with a barrier)
        if (tid < 8)
          snode[tid] = (ex + (tid & 1)) + nfnx * ((ey + ((tid >> 1) & 1)) + nfny * (ez + (tid >> 2)));
        for (int idx = tid; idx < 6 * 12; idx += 64)
          sg[idx / 12][idx % 12] = g[(int64_t)(idx % 12) * n_el + 6 * cell + idx / 12];
        if (tid < 6) sw[tid] = E[6 * cell + tid] * vol[6 * cell + tid];
        __syncthreads();
        if (tid < 24) {
          const int sidx = tid / 4, q = tid % 4;
          const int64_t node = conn[(int64_t)q * n_el + 6 * cell + sidx];
          int corner = -1;
          for (int c8 = 0; c8 < 8; ++c8)
            if (snode[c8] == node) corner = c8;
          scor[sidx][q] = corner;
        }
        if (tid < ND) sfix[tid] = fixed ? fixed[snode[tid / NC] * NC + tid % NC] : 0;
        for (int i = tid; i < NN; i += 64) sK[i] = 0.0;
        __syncthreads();
        // K_cell = sum_{s < 6} E_t V_t Z_t B^T D B Z_t: tet by tet (in order
        // s), each thread adding some of the tet's (4 NC)^2 entries; distinct
        // entries of a tet land on distinct cell entries
        constexpr int NL = 4 * NC;
        for (int sidx = 0; sidx < 6; ++sidx) {
          for (int e = tid; e < NL * NL; e += 64) {
            const int r = e / NL, col = e % NL;
            const int la = r / NC, ca = r % NC, lb = col / NC, cb = col % NC;
            const int qa = scor[sidx][la], qb = scor[sidx][lb];
            if (qa < 0 || qb < 0) continue;
            const int ia = qa * NC + ca, ib = qb * NC + cb;
            if (sfix[ia] || sfix[ib]) continue;
            const double* ga = sg[sidx] + 3 * la;
            const double* gb = sg[sidx] + 3 * lb;
            double v;
            if (NC == 1) {
              v = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
            } else {
              const double gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
              v = lam0 * ga[ca] * gb[cb] + mu0 * ((ca == cb ? gg : 0.0) + ga[cb] * gb[ca]);
            }
            sK[ia * ND + ib] += sw[sidx] * v;
          }
          __syncthreads();
        }
        galerkin_child<NC>(sK, sT, sP, a, b, c, fx, fy, fz, acc);
      }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = tid + 64 * k;
    if (i < NN) Kc[ce * NN + i] = acc[k];
  }
}

// On the finest level K_j = E_j K0 wherever the child has no Dirichlet DOF,
// so P_j^T K_j P_j = E_j M_j with M_j = P_j^T K0 P_j fixed per child
// position j = a + fx (b + fy c).  One workgroup per position.
template <int NC>
__global__ void __launch_bounds__(64)
k_galerkin_children(int fx, int fy, int fz, const double* __restrict__ K0b,
                    double* __restrict__ M) {
  constexpr int ND = 8 * NC;
  constexpr int NN = ND * ND;
  constexpr int PER = (NN + 63) / 64;
  __shared__ double sK[NN];
  __shared__ double sT[NN];
  __shared__ double sP[64];
  const int j = blockIdx.x;
  const int a = j % fx, b = (j / fx) % fy, c = j / (fx * fy);
  const int tid = threadIdx.x;
  for (int i = tid; i < NN; i += 64) sK[i] = K0b[i];
  double acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = 0.0;
  galerkin_child<NC>(sK, sT, sP, a, b, c, fx, fy, fz, acc);
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = tid + 64 * k;
    if (i < NN) M[(int64_t)j * NN + i] = acc[k];
  }
}

// 1 for the coarse elements of the first coarse level whose fine nodes hold
// a Dirichlet DOF (their children need the masked product), else 0.
template <int NC>
__global__ void k_galerkin_flags(int64_t ncx, int64_t ncy, int64_t ncz, int fx, int fy, int fz,
                                 const uint8_t* __restrict__ fixed, uint8_t* __restrict__ flag) {
  const int64_t nce = ncx * ncy * ncz;
  const int64_t nfnx = ncx * fx + 1, nfny = ncy * fy + 1;
  for (int64_t ce = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; ce < nce;
       ce += (int64_t)gridDim.x * blockDim.x) {
    const int64_t CX = ce % ncx, CY = (ce / ncx) % ncy, CZ = ce / (ncx * ncy);
    uint8_t any = 0;
    for (int k = 0; k <= fz; ++k)
      for (int j = 0; j <= fy; ++j)
        for (int i = 0; i <= fx; ++i) {
          const int64_t n = (CX * fx + i) + nfnx * ((CY * fy + j) + nfny * (CZ * fz + k));
#pragma unroll
          for (int d = 0; d < NC; ++d) any |= fixed[n * NC + d];
        }
    flag[ce] = any ? 1 : 0;
  }
}

// Kc = sum_j E_j M_j for the coarse elements without Dirichlet DOFs (flag
// null: all of them).  A workgroup of NN threads (one per entry) walks over
// coarse elements, with the element's child moduli staged in LDS; the
// children are summed in the order k_galerkin takes them.
constexpr int kMaxChildren = 125;  // factor 5 on every axis
template <int NC>
__global__ void __launch_bounds__(64 * NC * NC)
k_galerkin_interior(int64_t ncx, int64_t ncy, int64_t ncz, int fx, int fy, int fz,
                    const double* __restrict__ M, const double* __restrict__ E,
                    const uint8_t* __restrict__ flag, double* __restrict__ Kc) {
  constexpr int NN = 64 * NC * NC;
  __shared__ double sE[kMaxChildren];
  const int64_t nce = ncx * ncy * ncz;
  const int64_t nfx = ncx * fx, nfy = ncy * fy;
  const int nch = fx * fy * fz;
  const int tid = threadIdx.x;
  for (int64_t ce = blockIdx.x; ce < nce; ce += gridDim.x) {
    if (flag && flag[ce]) continue;  // uniform over the workgroup
    const int64_t CX = ce % ncx, CY = (ce / ncx) % ncy, CZ = ce / (ncx * ncy);
    for (int j = tid; j < nch; j += NN) {
      const int a = j % fx, b = (j / fx) % fy, c = j / (fx * fy);
      sE[j] = E[(CX * fx + a) + nfx * ((CY * fy + b) + nfy * (CZ * fz + c))];
    }
    __syncthreads();
    double acc = 0.0;
    for (int j = 0; j < nch; ++j) acc += sE[j] * M[(int64_t)j * NN + tid];
    Kc[ce * NN + tid] = acc;
    __syncthreads();
  }
}

// Sum of the element-matrix entries coupling (node n, comp d) with (node
// n + o, comp e) over the elements that hold both, in a fixed order.
template <int NC>
__device__ __forceinline__ double couple(const double* __restrict__ Kel, int64_t nx, int64_t ny,
                                         int64_t nz, int64_t i, int64_t j, int64_t k, int ox,
                                         int oy, int oz, int d, int e) {
  constexpr int ND = 8 * NC;
  double v = 0.0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int sx = s & 1, sy = (s >> 1) & 1, sz = s >> 2;
    const int64_t ex = i - 1 + sx, ey = j - 1 + sy, ez = k - 1 + sz;
    if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
    const int ax = 1 - sx, ay = 1 - sy, az = 1 - sz;
    const int bx = ax + ox, by = ay + oy, bz = az + oz;
    if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
    const int A = ax | (ay << 1) | (az << 2), B = bx | (by << 1) | (bz << 2);
    v += Kel[(ex + nx * (ey + ny * ez)) * (ND * ND) + (A * NC + d) * ND + B * NC + e];
  }
  return v;
}

// Stencil storage.  Coefficient c = (o * NC + d) * NC + e of a node couples
// its component d to component e of the neighbour at offset
// o = (ox+1) + 3 (oy+1) + 9 (oz+1).
// * Node-major: st[c * nn + node].
// * Tile-major (levels applied by the LDS-tiled kernel): nodes are grouped
//   in tsx x tsy x tsz tiles (tile-major, x fastest in both) and
//   st[(tile * 27 NC^2 + c) * ts + local], ts = tsx tsy tsz, padding nodes
//   past the grid holding zeros.  A wavefront's nodes then read contiguous
//   bytes per coefficient instead of 64-byte pieces of separate node rows;
//   the 60x30x30 level's apply went from 28.6 to 13.9 us.
// Tile shapes: 16x4x4 for levels of at least kTiledNodes nodes, 16x2x2
// (more, smaller workgroups) below: the 30x15x15 level's apply takes 6.9 us
// with it, 11.4 us with 16x4x4 tiles and 17.5 us node-major.
struct TileShape {
  int x, y, z;
};
constexpr TileShape kTileBig = {16, 4, 4}, kTileSmall = {16, 2, 2};
enum { ST_NODE = 0, ST_TILE_BIG = 1, ST_TILE_SMALL = 2 };

__host__ __device__ inline TileShape tile_shape(int layout) {
  return layout == ST_TILE_BIG ? kTileBig : kTileSmall;
}

__host__ __device__ inline int64_t st_tiles(int64_t nx, int64_t ny, int64_t nz, TileShape t) {
  return ((nx + t.x) / t.x) * ((ny + t.y) / t.y) * ((nz + t.z) / t.z);
}

// nodes the storage holds (tile padding included)
__host__ __device__ inline int64_t st_nodes(int64_t nx, int64_t ny, int64_t nz, int layout) {
  if (layout == ST_NODE) return (nx + 1) * (ny + 1) * (nz + 1);
  const TileShape t = tile_shape(layout);
  return st_tiles(nx, ny, nz, t) * (t.x * t.y * t.z);
}

// offset of coefficient c of node (i, j, k)
template <int NC>
__device__ __forceinline__ int64_t st_index(int64_t nx, int64_t ny, int64_t nz, int layout,
                                            int64_t i, int64_t j, int64_t k, int c) {
  const int64_t nnx = nx + 1, nny = ny + 1;
  if (layout != ST_NODE) {
    const TileShape t = tile_shape(layout);
    const int ts = t.x * t.y * t.z;
    const int64_t bx = (nx + t.x) / t.x, by = (ny + t.y) / t.y;
    const int64_t tile = i / t.x + bx * (j / t.y + by * (k / t.z));
    const int local = (int)(i % t.x + t.x * (j % t.y + t.y * (k % t.z)));
    return (tile * (27 * NC * NC) + c) * ts + local;
  }
  return c * (nnx * nny * (nz + 1)) + i + nnx * (j + nny * k);
}

// One entry per thread over the whole storage (tile padding included).
template <typename T, int NC>
__global__ void k_stencil_asm(int64_t nx, int64_t ny, int64_t nz, int layout,
                              const double* __restrict__ Kel, T* __restrict__ st) {
  constexpr int NS = 27 * NC * NC;
  const int64_t nnx = nx + 1, nny = ny + 1, nnz = nz + 1;
  const int64_t nn = nnx * nny * nnz;
  const int64_t total = st_nodes(nx, ny, nz, layout) * NS;
  const TileShape ts = tile_shape(layout);
  const int tsn = ts.x * ts.y * ts.z;
  const int64_t bx = (nx + ts.x) / ts.x, by = (ny + ts.y) / ts.y;
  // Consecutive blocks of entries (consecutive components of a tile) read
  // the same element-matrix lines; workgroups go to the 8 XCDs (each with
  // its own L2) in turn, so XCD x takes the contiguous range of logical
  // blocks starting at x q + min(x, r) (q, r: quotient and remainder of the
  // grid by 8) and those reads meet in one L2.
  const int64_t nb = gridDim.x, q8 = nb / 8, r8 = nb % 8;
  const int64_t xcd = blockIdx.x % 8;
  const int64_t lb = xcd * q8 + (xcd < r8 ? xcd : r8) + blockIdx.x / 8;
  for (int64_t t = lb * blockDim.x + threadIdx.x; t < total; t += nb * blockDim.x) {
    int64_t i, j, k;
    int c;
    if (layout != ST_NODE) {
      const int local = (int)(t % tsn);
      c = (int)((t / tsn) % NS);
      const int64_t tile = t / ((int64_t)tsn * NS);
      i = (tile % bx) * ts.x + local % ts.x;
      j = ((tile / bx) % by) * ts.y + (local / ts.x) % ts.y;
      k = (tile / (bx * by)) * ts.z + local / (ts.x * ts.y);
      if (i >= nnx || j >= nny || k >= nnz) {
        st[t] = T(0);
        continue;
      }
    } else {
      const int64_t node = t % nn;
      c = (int)(t / nn);
      i = node % nnx;
      j = (node / nnx) % nny;
      k = node / (nnx * nny);
    }
    const int e = c % NC, d = (c / NC) % NC, o = c / (NC * NC);
    const int ox = o % 3 - 1, oy = (o / 3) % 3 - 1, oz = o / 9 - 1;
    st[t] = (T)couple<NC>(Kel, nx, ny, nz, i, j, k, ox, oy, oz, d, e);
  }
}

// Jacobi inverse of a stencil level; rows with no stiffness (DOFs that no
// free fine DOF depends on) get 0 and stay 0 through the cycle.
template <typename T, int NC>
__global__ void k_stencil_dinv(int64_t nx, int64_t ny, int64_t nz, int layout,
                               const T* __restrict__ st, T* __restrict__ dinv) {
  const int64_t nnx = nx + 1, nny = ny + 1, nn = nnx * nny * (nz + 1);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn * NC;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = t / NC;
    const int d = (int)(t % NC);
    const int64_t i = node % nnx, j = (node / nnx) % nny, k = node / (nnx * nny);
    const T v = st[st_index<NC>(nx, ny, nz, layout, i, j, k, (13 * NC + d) * NC + d)];
    dinv[t] = v > T(0) ? T(1) / v : T(0);
  }
}

// Dense coarsest matrix (fp64, n x n row-major); zero rows become identity.
template <int NC>
__global__ void k_dense_asm(int64_t nx, int64_t ny, int64_t nz, const double* __restrict__ Kel,
                            double* __restrict__ A) {
  const int64_t nnx = nx + 1, nny = ny + 1;
  const int64_t n = nnx * nny * (nz + 1) * NC;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / n, c = t % n;
    const int64_t nr = r / NC, ncn = c / NC;
    const int d = (int)(r % NC), e = (int)(c % NC);
    const int64_t i = nr % nnx, j = (nr / nnx) % nny, k = nr / (nnx * nny);
    const int64_t i2 = ncn % nnx, j2 = (ncn / nnx) % nny, k2 = ncn / (nnx * nny);
    const int64_t ox = i2 - i, oy = j2 - j, oz = k2 - k;
    double v = 0.0;
    if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1)
      v = couple<NC>(Kel, nx, ny, nz, i, j, k, (int)ox, (int)oy, (int)oz, d, e);
    A[t] = v;
  }
}

__global__ void k_dense_fix_diag(int64_t n, double* __restrict__ A) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x)
    if (!(A[r * n + r] > 0.0)) A[r * n + r] = 1.0;
}

// ------------------------------------------- blocked Gauss-Jordan inverse
// In-place inversion of an SPD matrix (no pivoting needed) by pivot blocks B
// of kGJ rows: with R the remaining indices,
//   A_BB <- A_BB^-1, A_BR <- A_BB^-1 A_BR, A_RR <- A_RR - A_RB A_BR(new),
//   A_RB <- -A_RB A_BB^-1.
// The three off-block updates are LDS-tiled: kTile x kTile tiles for A_RR
// (the block dimension in two halves), kTileRC-wide strips for A_BR, A_RB.
constexpr int kGJ = 64;
constexpr int kTile = 64;
constexpr int kTileRC = 32;

// invert the pivot block (padded with identity), write it to Binv
// (kGJ x kGJ) and back into A.  Each thread keeps a 4 x 4 block of the
// matrix in registers; per pivot k the owners of row k and column k publish
// them through LDS (double-buffered, so one barrier per pivot), and every
// thread applies a_ij -= a_ik a_kj / a_kk with column k zeroed first and row
// k replaced afterwards:
//   row k:            a_kj / a_kk (j != k),  1 / a_kk (j = k)
//   column k, i != k: -a_ik / a_kk
// The pivot index inside a thread's 4 x 4 block is a compile-time constant
// of the unrolled inner loop, so there is no per-element branching.
__device__ __forceinline__ void gj_block_inv(int64_t n, int64_t b0, int nb,
                                             double* __restrict__ A,
                                             double* __restrict__ Binv) {
  static_assert(kGJ == 64, "4 x 4 register blocks of a 64 x 64 matrix");
  __shared__ double prow[2][kGJ], pcol[2][kGJ];
  const int tid = threadIdx.x;
  const int tr = tid / 16, tc = tid % 16;  // rows 4 tr.., columns 4 tc..
  double m[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 4 * tr + i, c = 4 * tc + j;
      m[i][j] = (r < nb && c < nb) ? A[(b0 + r) * n + b0 + c] : (r == c ? 1.0 : 0.0);
    }
  for (int kb = 0; kb < kGJ / 4; ++kb) {
#pragma unroll
    for (int ki = 0; ki < 4; ++ki) {
      const int k = 4 * kb + ki, buf = ki & 1;
      const bool own_row = tr == kb, own_col = tc == kb;
      if (own_row) {
#pragma unroll
        for (int j = 0; j < 4; ++j) prow[buf][4 * tc + j] = m[ki][j];
      }
      if (own_col) {
#pragma unroll
        for (int i = 0; i < 4; ++i) pcol[buf][4 * tr + i] = m[i][ki];
      }
      __syncthreads();
      const double ip = 1.0 / pcol[buf][k];
      double rk[4], ck[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) rk[j] = prow[buf][4 * tc + j] * ip;
#pragma unroll
      for (int i = 0; i < 4; ++i) ck[i] = pcol[buf][4 * tr + i];
      if (own_col) {
        rk[ki] = ip;
#pragma unroll
        for (int i = 0; i < 4; ++i) m[i][ki] = 0.0;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) m[i][j] -= ck[i] * rk[j];
      if (own_row) {
#pragma unroll
        for (int j = 0; j < 4; ++j) m[ki][j] = rk[j];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 4 * tr + i, c = 4 * tc + j;
      Binv[r * kGJ + c] = m[i][j];
      if (r < nb && c < nb) A[(b0 + r) * n + b0 + c] = m[i][j];
    }
}

__global__ void __launch_bounds__(256) k_gj_block_inv(int64_t n, int64_t b0, int nb,
                                                       double* __restrict__ A,
                                                       double* __restrict__ Binv) {
  gj_block_inv(n, b0, nb, A, Binv);
}

// A_BR <- Binv A_BR on a strip of kTileRC columns (columns inside the block
// are left alone).  Binv is the inverse padded with identity and the strip
// rows past nb are zero, so the products run over all kGJ rows; each thread
// keeps 8 outputs (one column, rows tid / 32 + 8 m) in flight.
__global__ void __launch_bounds__(256) k_gj_rows(int64_t n, int64_t b0, int nb,
                                                  const double* __restrict__ Binv,
                                                  double* __restrict__ A) {
  static_assert(kGJ * kTileRC == 8 * 256, "8 outputs per thread");
  __shared__ double sB[kGJ][kGJ + 1];
  __shared__ double sX[kGJ][kTileRC + 1];
  const int tid = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * kTileRC;
  for (int i = tid; i < kGJ * kGJ; i += 256) sB[i / kGJ][i % kGJ] = Binv[i];
  for (int i = tid; i < kGJ * kTileRC; i += 256) {
    const int r = i / kTileRC, c = i % kTileRC;
    sX[r][c] = (r < nb && c0 + c < n) ? A[(b0 + r) * n + c0 + c] : 0.0;
  }
  __syncthreads();
  const int c = tid % kTileRC, rb = tid / kTileRC;
  double acc[8] = {};
#pragma unroll 8
  for (int q = 0; q < kGJ; ++q) {
    const double x = sX[q][c];
#pragma unroll
    for (int mm = 0; mm < 8; ++mm) acc[mm] += sB[rb + 8 * mm][q] * x;
  }
  const int64_t col = c0 + c;
  if (col >= n || (col >= b0 && col < b0 + nb)) return;
#pragma unroll
  for (int mm = 0; mm < 8; ++mm) {
    const int r = rb + 8 * mm;
    if (r < nb) A[(b0 + r) * n + col] = acc[mm];
  }
}

// A_RR <- A_RR - A_RB A_BR on a kTile x kTile tile (rows or columns inside
// the block are left alone)
__global__ void __launch_bounds__(256) k_gj_update(int64_t n, int64_t b0, int nb,
                                                    double* __restrict__ A) {
  constexpr int kH = kGJ / 2;             // block dimension in two halves
  __shared__ double sL[kTile][kH + 1];    // A_RB rows of the tile
  __shared__ double sU[kH][kTile + 1];    // A_BR columns of the tile
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * kTile, c0 = (int64_t)blockIdx.x * kTile;
  // 4 x 4 outputs per thread: rows tr + 16 a, columns tc + 16 b
  const int tr = tid / 16, tc = tid % 16;
  double acc[4][4] = {};
  for (int h = 0; h < kGJ; h += kH) {
    for (int i = tid; i < kTile * kH; i += 256) {
      const int r = i / kH, q = i % kH;
      sL[r][q] = (r0 + r < n && h + q < nb) ? A[(r0 + r) * n + b0 + h + q] : 0.0;
    }
    for (int i = tid; i < kH * kTile; i += 256) {
      const int q = i / kTile, c = i % kTile;
      sU[q][c] = (h + q < nb && c0 + c < n) ? A[(b0 + h + q) * n + c0 + c] : 0.0;
    }
    __syncthreads();
    for (int q = 0; q < kH; ++q) {
      double l[4], u[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) l[a] = sL[tr + 16 * a][q];
#pragma unroll
      for (int b = 0; b < 4; ++b) u[b] = sU[q][tc + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += l[a] * u[b];
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t r = r0 + tr + 16 * a;
    if (r >= n || (r >= b0 && r < b0 + nb)) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t c = c0 + tc + 16 * b;
      if (c >= n || (c >= b0 && c < b0 + nb)) continue;
      A[r * n + c] -= acc[a][b];
    }
  }
}

// A_RB <- -A_RB Binv on a strip of kTileRC rows (rows inside the block are
// left alone); each thread keeps 8 outputs (one column, rows tid / 64 + 4 m)
// in flight, the products over all kGJ columns as in k_gj_rows.
__device__ __forceinline__ void gj_cols(int64_t n, int64_t b0, int nb, int strip,
                                        const double* __restrict__ Binv,
                                        double* __restrict__ A) {
  static_assert(kGJ * kTileRC == 8 * 256, "8 outputs per thread");
  __shared__ double sB[kGJ][kGJ + 1];
  __shared__ double sX[kTileRC][kGJ + 1];
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)strip * kTileRC;
  for (int i = tid; i < kGJ * kGJ; i += 256) sB[i / kGJ][i % kGJ] = Binv[i];
  for (int i = tid; i < kTileRC * kGJ; i += 256) {
    const int r = i / kGJ, q = i % kGJ;
    sX[r][q] = (r0 + r < n && q < nb) ? A[(r0 + r) * n + b0 + q] : 0.0;
  }
  __syncthreads();
  const int c = tid % kGJ, rb = tid / kGJ;
  double acc[8] = {};
#pragma unroll 8
  for (int q = 0; q < kGJ; ++q) {
    const double b = sB[q][c];
#pragma unroll
    for (int mm = 0; mm < 8; ++mm) acc[mm] += sX[rb + 4 * mm][q] * b;
  }
  if (c >= nb) return;
#pragma unroll
  for (int mm = 0; mm < 8; ++mm) {
    const int64_t row = r0 + rb + 4 * mm;
    if (row >= n || (row >= b0 && row < b0 + nb)) continue;
    A[row * n + b0 + c] = -acc[mm];
  }
}

// Column strips of pivot block b0 (Binv) and, in the one extra workgroup,
// the inversion of the next pivot block b1 (nb1 > 0) into Binv1: that block
// is final once the trailing update of b0 has run, and neither touches the
// other's entries or inverse buffer.
__global__ void __launch_bounds__(256) k_gj_cols_next(int64_t n, int64_t b0, int nb,
                                                       const double* __restrict__ Binv,
                                                       int64_t b1, int nb1,
                                                       double* __restrict__ Binv1,
                                                       double* __restrict__ A) {
  if (blockIdx.x + 1 == gridDim.x) {
    if (nb1 > 0) gj_block_inv(n, b1, nb1, A, Binv1);
    return;
  }
  gj_cols(n, b0, nb, blockIdx.x, Binv, A);
}

template <typename T>
__global__ void k_cast(int64_t n, const double* __restrict__ a, T* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (T)a[i];
}

// ------------------------------------------------------------ operators

// N consecutive values of a node (NC components x KC columns, contiguous in
// the blocked layout) with paired loads / stores when KC is even: the
// offset is then a multiple of 2 values, so 8-byte (float) or 16-byte
// (double) aligned.
template <typename T> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };

template <typename T, int N, int KC>
__device__ __forceinline__ void load_node(const T* __restrict__ p, T (&v)[N]) {
  if constexpr (KC % 2 == 0) {
    using V = typename Pair<T>::type;
    const V* q = reinterpret_cast<const V*>(p);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
      const V w = q[i];
      v[2 * i] = w.x;
      v[2 * i + 1] = w.y;
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = p[i];
  }
}

template <typename T, int N, int KC>
__device__ __forceinline__ void store_node(T* __restrict__ p, const T (&v)[N]) {
  if constexpr (KC % 2 == 0) {
    using V = typename Pair<T>::type;
    V* q = reinterpret_cast<V*>(p);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
      V w;
      w.x = v[2 * i];
      w.y = v[2 * i + 1];
      q[i] = w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) p[i] = v[i];
  }
}

// Output of a stencil apply: mode 0 writes y = S x; mode 2 writes a fused
// Chebyshev-Jacobi step x_out = x + c1 (x - x_prev) + c2 dinv (b - S x),
// where x_prev is read from the output buffer itself (ping-pong iterates,
// so no direction vector exists) unless prev_zero, and (c1, c2) = coef[0..1].
enum { ST_APPLY = 0, ST_CHEB = 2 };

template <typename T>
struct StOut {
  const T* b = nullptr;
  const T* dinv = nullptr;
  const double* coef = nullptr;
  int prev_zero = 1;
  const double* stop = nullptr;  // column stop flags (tiled kernel: skip stopped workgroups)
};

// y = S x on a stencil level.  A thread takes one (node, component row) and
// up to GM column groups, so each stencil row is read once for all of them.
template <typename T, int NC, int KC, int GM, int MODE>
__global__ void k_stencil_apply(int64_t nx, int64_t ny, int64_t nz, int groups,
                                const T* __restrict__ st, const T* __restrict__ X,
                                T* __restrict__ Y, StOut<T> oa) {
  const int64_t nnx = nx + 1, nny = ny + 1, nnz = nz + 1;
  const int64_t nn = nnx * nny * nnz;
  const int chunks = (groups + GM - 1) / GM;
  const int64_t total = nn * NC * chunks;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = t % nn;
    const int d = (int)((t / nn) % NC);
    const int g0 = (int)(t / (nn * NC)) * GM;
    const int ng = groups - g0 < GM ? groups - g0 : GM;
    const int64_t i = node % nnx, j = (node / nnx) % nny, k = node / (nnx * nny);
    T acc[GM][KC];
#pragma unroll
    for (int q = 0; q < GM; ++q)
#pragma unroll
      for (int c = 0; c < KC; ++c) acc[q][c] = T(0);
    // unrolled and branch-free: a neighbour outside the grid has an exactly
    // zero stencil block, so its address is clamped to the node itself
#pragma unroll
    for (int o = 0; o < 27; ++o) {
      const int ox = o % 3 - 1, oy = (o / 3) % 3 - 1, oz = o / 9 - 1;
      const int64_t i2 = i + ox, j2 = j + oy, k2 = k + oz;
      const bool in = i2 >= 0 && i2 < nnx && j2 >= 0 && j2 < nny && k2 >= 0 && k2 < nnz;
      const int64_t m = in ? i2 + nnx * (j2 + nny * k2) : node;
      T a[NC];
#pragma unroll
      for (int e = 0; e < NC; ++e) a[e] = st[((o * NC + d) * NC + e) * nn + node];
#pragma unroll
      for (int q = 0; q < GM; ++q) {
        if (q >= ng) break;
        T xv[NC * KC];
        load_node<T, NC * KC, KC>(X + ((int64_t)(g0 + q) * nn + m) * NC * KC, xv);
#pragma unroll
        for (int e = 0; e < NC; ++e)
#pragma unroll
          for (int c = 0; c < KC; ++c) acc[q][c] += a[e] * xv[e * KC + c];
      }
    }
    T c1 = T(0), c2 = T(0), dv = T(0);
    if constexpr (MODE == ST_CHEB) {
      c1 = (T)oa.coef[0];
      c2 = (T)oa.coef[1];
      dv = oa.dinv[node * NC + d];
    }
#pragma unroll
    for (int q = 0; q < GM; ++q) {
      if (q >= ng) break;
      const int64_t off = (((int64_t)(g0 + q) * nn + node) * NC + d) * KC;
      if constexpr (MODE == ST_CHEB) {
        T xc[KC], bv[KC], xp[KC];
        load_node<T, KC, KC>(X + off, xc);
        load_node<T, KC, KC>(oa.b + off, bv);
        if (oa.prev_zero) {
#pragma unroll
          for (int c = 0; c < KC; ++c) xp[c] = T(0);
        } else {
          load_node<T, KC, KC>(Y + off, xp);
        }
#pragma unroll
        for (int c = 0; c < KC; ++c)
          acc[q][c] = xc[c] + c1 * (xc[c] - xp[c]) + c2 * dv * (bv[c] - acc[q][c]);
      }
      store_node<T, KC, KC>(Y + off, acc[q]);
    }
  }
}

// First Chebyshev step from a zero iterate on the fine level:
// x = c2 dinv b (blocked vectors, total = ndof * k)
template <typename T>
__global__ void k_cheb_first(int64_t total, int64_t ndof, int kc, const double* __restrict__ coef,
                             const T* __restrict__ b, const T* __restrict__ dinv,
                             T* __restrict__ x) {
  const T c2 = (T)coef[1];
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x)
    x[t] = c2 * dinv[(t / kc) % ndof] * b[t];
}

// y = S x on a stencil level, LDS-tiled: a workgroup owns a SX x SY x SZ
// block of nodes and stages x of the block and its one-node halo for GM
// column groups in LDS once.  NC threads share a node, one per component
// row, so a workgroup runs TS NC threads (TS = SX SY SZ): each sums its
// row's 27 stencil blocks for all GM groups.  Stencil rows are read once per
// node, from the tile-major storage of the same tile shape.
template <typename T, int NC, int KC, int GM, int MODE, int SX, int SY, int SZ>
__global__ void __launch_bounds__(SX * SY * SZ * NC)
k_stencil_apply_tiled(int64_t nx, int64_t ny, int64_t nz, int groups, const T* __restrict__ st,
                      const T* __restrict__ X, T* __restrict__ Y, StOut<T> oa) {
  constexpr int HX = SX + 2, HY = SY + 2, HZ = SZ + 2, TS = SX * SY * SZ;
  constexpr int V = NC * KC;          // values of one node and one group
  constexpr int W = V * GM;           // values of one node in LDS
  __shared__ T xs[HZ * HY * HX * W];
  const int64_t nnx = nx + 1, nny = ny + 1, nnz = nz + 1;
  const int64_t nn = nnx * nny * nnz;
  const int bx = (int)((nnx + SX - 1) / SX), by = (int)((nny + SY - 1) / SY),
            bz = (int)((nnz + SZ - 1) / SZ);
  int b = blockIdx.x;
  const int tbx = b % bx; b /= bx;
  const int tby = b % by; b /= by;
  const int tbz = b % bz; b /= bz;
  const int g0 = b * GM;
  const int ng = groups - g0 < GM ? groups - g0 : GM;
  if (oa.stop) {
    bool all = true;
    for (int c = g0 * KC; c < (g0 + ng) * KC; ++c) all = all && oa.stop[c * kStopStride] != 0.0;
    if (all) return;
  }
  const int64_t x0 = (int64_t)tbx * SX, y0 = (int64_t)tby * SY, z0 = (int64_t)tbz * SZ;
  // this tile's coefficients (tile-major storage)
  st += (int64_t)(tbx + bx * (tby + by * tbz)) * (27 * NC * NC) * TS;
  const int nt = threadIdx.x % TS, d = threadIdx.x / TS;
  // stage x (halo nodes outside the grid read as zero; their stencil
  // entries are zero as well).  A thread's NIT items are all loaded before
  // any is stored, with unpredicated loads (an item outside the grid reads
  // the first node of the vector and stores zero): a load in a branch of its
  // own is waited for inside it, one memory latency per item.
  constexpr int NH = HZ * HY * HX * GM, NIT = (NH + TS * NC - 1) / (TS * NC);
  {
    T v[NIT][V];
    bool ok[NIT];
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
      const int it = threadIdx.x + u * TS * NC;
      const int q = it % GM, hn = it / GM;
      const int hx = hn % HX, hy = (hn / HX) % HY, hz = hn / (HX * HY);
      const int64_t gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
      ok[u] = it < NH && q < ng && gx >= 0 && gx < nnx && gy >= 0 && gy < nny && gz >= 0 &&
              gz < nnz;
      const int64_t off = ok[u] ? ((int64_t)(g0 + q) * nn + gx + nnx * (gy + nny * gz)) * V : 0;
      load_node<T, V, KC>(X + off, v[u]);
    }
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
      const int it = threadIdx.x + u * TS * NC;
      if (it < NH) {
        const int q = it % GM, hn = it / GM;
#pragma unroll
        for (int i = 0; i < V; ++i) xs[hn * W + q * V + i] = ok[u] ? v[u][i] : T(0);
      }
    }
  }
  __syncthreads();
  const int tx = nt % SX, ty = (nt / SX) % SY, tz = nt / (SX * SY);
  const int64_t gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
  if (gx >= nnx || gy >= nny || gz >= nnz) return;
  const int64_t node = gx + nnx * (gy + nny * gz);
  T acc[GM][KC];
#pragma unroll
  for (int q = 0; q < GM; ++q)
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[q][c] = T(0);
#pragma unroll 3
  for (int o = 0; o < 27; ++o) {
    const int ox = o % 3, oy = (o / 3) % 3, oz = o / 9;  // halo offsets 0..2
    const T* xn = xs + ((tz + oz) * HY * HX + (ty + oy) * HX + (tx + ox)) * W;
    T a[NC];
#pragma unroll
    for (int e = 0; e < NC; ++e) a[e] = st[((o * NC + d) * NC + e) * TS + nt];
#pragma unroll
    for (int q = 0; q < GM; ++q)
#pragma unroll
      for (int e = 0; e < NC; ++e)
#pragma unroll
        for (int c = 0; c < KC; ++c) acc[q][c] += a[e] * xn[q * V + e * KC + c];
  }
  const T* xcen = xs + ((tz + 1) * HY * HX + (ty + 1) * HX + (tx + 1)) * W;
  // The Chebyshev operands of every group are loaded together (unpredicated:
  // a group past the last, or x_prev when it is zero, reads offset 0).  With
  // every group present (nq = GM, a constant after inlining) no store is
  // predicated, so the compiler does not sink a group's loads into a branch
  // of its own, where each would be waited for in turn.
  auto finish = [&](const int nq) {
    T c1 = T(0), c2 = T(0), dv = T(0);
    T bv[GM][KC], xp[GM][KC];
    if constexpr (MODE == ST_CHEB) {
      c1 = (T)oa.coef[0];
      c2 = (T)oa.coef[1];
      dv = oa.dinv[node * NC + d];
#pragma unroll
      for (int q = 0; q < GM; ++q) {
        const int64_t off = ((int64_t)(g0 + (q < nq ? q : 0)) * nn + node) * V + d * KC;
        load_node<T, KC, KC>(oa.b + off, bv[q]);
        load_node<T, KC, KC>(Y + (oa.prev_zero ? 0 : off), xp[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < GM; ++q) {
      if (q >= nq) break;
      const int64_t off = ((int64_t)(g0 + q) * nn + node) * V + d * KC;
      if constexpr (MODE == ST_CHEB) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const T xc = xcen[q * V + d * KC + c];
          const T xpc = oa.prev_zero ? T(0) : xp[q][c];
          acc[q][c] = xc + c1 * (xc - xpc) + c2 * dv * (bv[q][c] - acc[q][c]);
        }
      }
      store_node<T, KC, KC>(Y + off, acc[q]);
    }
  };
  if (ng == GM)
    finish(GM);
  else
    finish(ng);
}

// Chebyshev step on a level (blocked vectors, total = ndof * k):
//   d <- c1 d + c2 dinv (b - w);  x <- x + d   (w null: the iterate is 0)
// coef holds (c1, c2) for this step.
template <typename T>
__global__ void k_cheb(int64_t total, int64_t ndof, int kc, const double* __restrict__ coef,
                       const T* __restrict__ b, const T* __restrict__ w,
                       const T* __restrict__ dinv, T* __restrict__ d, T* __restrict__ x,
                       int x_is_zero) {
  const T c1 = (T)coef[0], c2 = (T)coef[1];
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t dof = (t / kc) % ndof;
    const T r = w ? b[t] - w[t] : b[t];
    const T dn = ((x_is_zero || c1 == T(0)) ? T(0) : c1 * d[t]) + c2 * dinv[dof] * r;
    d[t] = dn;
    x[t] = x_is_zero ? dn : x[t] + dn;
  }
}

// w = b - w (the fine level's residual after a generic operator apply), zero at
// Dirichlet DOFs
template <typename T>
__global__ void k_rsub(int64_t total, int64_t ndof, int kc, const uint8_t* __restrict__ fixed,
                       const T* __restrict__ b, T* __restrict__ w) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x)
    w[t] = (fixed && fixed[(t / kc) % ndof]) ? T(0) : b[t] - w[t];
}

template <typename T>
__global__ void k_copy_vec(int64_t total, const T* __restrict__ a, T* __restrict__ b) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x)
    b[t] = a[t];
}

// Transfers run on row grids: blockIdx.x and threadIdx.x give the node along
// x, blockIdx.y (strided) the row (y, z, column group), so the index
// arithmetic is a few 32-bit divisions per row rather than 64-bit ones per
// value.  (Splitting a node's components over three lanes made the
// prolongation twice as slow: the row of 3 nx pieces fills 256-lane
// workgroups poorly and every lane repeats the weights.)  Row counts and row
// lengths fit in 32 bits.
struct RowGrid {
  dim3 grid, block;
};

static RowGrid row_grid(int64_t row_len, int64_t rows) {
  RowGrid g;
  const int bx = row_len >= 256 ? 256 : (int)((row_len + 63) / 64) * 64;
  g.block = dim3((unsigned)bx);
  g.grid = dim3((unsigned)((row_len + bx - 1) / bx), (unsigned)std::min<int64_t>(rows, 65535));
  return g;
}

// Restriction b_c = P^T Z (b - w), separable in two passes.
// Pass 1 (x, y): t[(I, J, k)] = sum_{dx,dy} w(dx) w(dy) r(f I + dx, f J + dy, k)
// on every fine z plane k, r = b - w with Dirichlet DOFs removed.  Thread:
// coarse I; row: (J, k, group).
template <typename T, int NC, int KC>
__global__ void k_restrict_xy(int fnx, int fny, int fnz, int fx, int fy, int groups,
                              const T* __restrict__ b, const T* __restrict__ w,
                              const uint8_t* __restrict__ fixed, T* __restrict__ tbuf) {
  const int fnnx = fnx + 1, fnny = fny + 1, fnnz = fnz + 1;
  const int cnnx = fnx / fx + 1, cnny = fny / fy + 1;
  const int64_t tn = (int64_t)cnnx * cnny * fnnz, fnn = (int64_t)fnnx * fnny * fnnz;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= cnnx) return;
  const int nrows = cnny * fnnz * groups;
  for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
    const int J = row % cnny, kg = row / cnny, k = kg % fnnz, g = kg / fnnz;
    const T* bg = b + g * fnn * NC * KC;
    const T* wg = w ? w + g * fnn * NC * KC : nullptr;
    T acc[NC][KC];
#pragma unroll
    for (int d = 0; d < NC; ++d)
#pragma unroll
      for (int c = 0; c < KC; ++c) acc[d][c] = T(0);
    for (int dy = -(fy - 1); dy <= fy - 1; ++dy) {
      const int j = J * fy + dy;
      if (j < 0 || j >= fnny) continue;
      const T wy = T(1) - (T)(dy < 0 ? -dy : dy) / (T)fy;
      for (int dx = -(fx - 1); dx <= fx - 1; ++dx) {
        const int i = I * fx + dx;
        if (i < 0 || i >= fnnx) continue;
        const T wgt = (T(1) - (T)(dx < 0 ? -dx : dx) / (T)fx) * wy;
        const int64_t fn = i + fnnx * (j + (int64_t)fnny * k);
        T bv[NC * KC], wv[NC * KC];
        load_node<T, NC * KC, KC>(bg + fn * NC * KC, bv);
        if (wg) load_node<T, NC * KC, KC>(wg + fn * NC * KC, wv);
#pragma unroll
        for (int d = 0; d < NC; ++d) {
          if (fixed && fixed[fn * NC + d]) continue;
#pragma unroll
          for (int c = 0; c < KC; ++c)
            acc[d][c] += wgt * (wg ? bv[d * KC + c] - wv[d * KC + c] : bv[d * KC + c]);
        }
      }
    }
    const int64_t tnode = I + cnnx * (J + (int64_t)cnny * k);
    T* out = tbuf + g * tn * NC * KC;
#pragma unroll
    for (int d = 0; d < NC; ++d)
#pragma unroll
      for (int c = 0; c < KC; ++c) out[(tnode * NC + d) * KC + c] = acc[d][c];
  }
}

// Pass 2 (z): b_c(I, J, K) = sum_dz w(dz) t(I, J, f K + dz).  Thread: (I, J);
// row: (K, group).
template <typename T, int NC, int KC>
__global__ void k_restrict_z(int cnnx, int cnny, int fnz, int fz, int groups,
                             const T* __restrict__ tbuf, T* __restrict__ bc) {
  const int fnnz = fnz + 1, cnnz = fnz / fz + 1;
  const int64_t plane = (int64_t)cnnx * cnny;
  const int64_t tn = plane * fnnz, cnn = plane * cnnz;
  const int64_t IJ = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (IJ >= plane) return;
  const int nrows = cnnz * groups;
  for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
    const int K = row % cnnz, g = row / cnnz;
    const T* tg = tbuf + g * tn * NC * KC;
    T acc[NC][KC];
#pragma unroll
    for (int d = 0; d < NC; ++d)
#pragma unroll
      for (int c = 0; c < KC; ++c) acc[d][c] = T(0);
    for (int dz = -(fz - 1); dz <= fz - 1; ++dz) {
      const int k = K * fz + dz;
      if (k < 0 || k >= fnnz) continue;
      const T wz = T(1) - (T)(dz < 0 ? -dz : dz) / (T)fz;
      const int64_t tnode = IJ + plane * k;
#pragma unroll
      for (int d = 0; d < NC; ++d)
#pragma unroll
        for (int c = 0; c < KC; ++c) acc[d][c] += wz * tg[(tnode * NC + d) * KC + c];
    }
    const int64_t cnode = IJ + plane * K;
    T* out = bc + g * cnn * NC * KC;
#pragma unroll
    for (int d = 0; d < NC; ++d)
#pragma unroll
      for (int c = 0; c < KC; ++c) out[(cnode * NC + d) * KC + c] = acc[d][c];
  }
}

// Restriction for factor 2 on every axis in one pass, b_c = P^T (b - w)
// (HASW: w given).  A 256-thread workgroup owns kR2I x kR2J coarse nodes of
// one column group and a chunk of zlen coarse planes (one where that leaves
// too few workgroups to hide the planes' latency), and streams the fine planes
// 2 K0 - 1 .. 2 K1 - 1 through LDS: per plane, the tile's (2 kR2I + 1) x
// (2 kR2J + 1) fine nodes, each fine row a contiguous run of the blocked
// layout (coalesced), prefetched one plane ahead in registers.  Each thread
// sums its coarse node's 3 x 3 fine neighbourhood on a plane with the
// weights and in the order of k_restrict_xy, and the planes with those of
// k_restrict_z, so the result is the two-pass one bit for bit.  No Dirichlet
// mask is applied: b must be zero at Dirichlet DOFs (the fine residual is) or
// there are none (coarse levels).  With x0 the kernel also writes the coarse
// level's first smoothing iterate from zero, x0 = c2 dinv_c b_c (k_cheb_first
// with c2 = cf[1]).
constexpr int kR2I = 32, kR2J = 8;

template <typename T, int NC, int KC, bool HASW>
__global__ void __launch_bounds__(256)
k_restrict222(int fnx, int fny, int fnz, int tiles_x, int tiles_y, int zchunks, int zlen,
              const T* __restrict__ b, const T* __restrict__ w, T* __restrict__ bc,
              const T* __restrict__ dinv_c, const double* __restrict__ cf, T* __restrict__ x0) {
  static_assert(kR2I * kR2J == 256, "one thread per coarse node of a tile");
  constexpr int NV = NC * KC;
  constexpr int kW = 2 * kR2I + 1, kH = 2 * kR2J + 1;  // staged fine nodes per row, rows
  constexpr int kRow = kW * NV;                          // values of one staged row
  constexpr int kM = (kRow + 63) / 64;                   // values per lane per row
  constexpr int kQ = (kH + 3) / 4;                       // rows per wavefront
  __shared__ T sv[NV * kH * kW];                         // [value][row][node]
  const int fnnx = fnx + 1, fnny = fny + 1, fnnz = fnz + 1;
  const int cnnx = fnx / 2 + 1, cnny = fny / 2 + 1, cnnz = fnz / 2 + 1;
  const int64_t fnn = (int64_t)fnnx * fnny * fnnz, cnn = (int64_t)cnnx * cnny * cnnz;
  int blk = blockIdx.x;
  const int tx = blk % tiles_x;
  blk /= tiles_x;
  const int ty = blk % tiles_y;
  blk /= tiles_y;
  const int zc = blk % zchunks;
  const int g = blk / zchunks;
  const int K0 = zc * zlen, K1 = K0 + zlen < cnnz ? K0 + zlen : cnnz;
  const int gx0 = 2 * kR2I * tx - 1, gy0 = 2 * kR2J * ty - 1;
  const T* bg = b + g * fnn * NV;
  const T* wg = HASW ? w + g * fnn * NV : nullptr;
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  // staging: wavefront wv takes rows wv, wv + 4, ...; lane l takes values
  // l, l + 64, ... of a row (value e: fine node gx0 + e / NV, component and
  // column e % NV)
  int s_lds[kM];
  unsigned s_ok = 0;
#pragma unroll
  for (int m = 0; m < kM; ++m) {
    const int e = lane + 64 * m;
    const int xn = e / NV, v = e % NV;
    s_lds[m] = (v * kH) * kW + xn;
    if (e < kRow && gx0 + xn >= 0 && gx0 + xn < fnnx) s_ok |= 1u << m;
  }
  T pb[kQ][kM], pw[HASW ? kQ : 1][HASW ? kM : 1];
  auto fetch = [&](int gz) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const int r = wv + 4 * q, gy = gy0 + r;
      const bool row_ok = r < kH && gy >= 0 && gy < fnny && gz >= 0 && gz < fnnz;
      const int64_t rowoff = (((int64_t)gz * fnny + gy) * fnnx + gx0) * NV;
#pragma unroll
      for (int m = 0; m < kM; ++m) {
        const bool ok = row_ok && ((s_ok >> m) & 1u);
        const int64_t o = rowoff + lane + 64 * m;
        pb[q][m] = ok ? bg[o] : T(0);
        if constexpr (HASW) pw[q][m] = ok ? wg[o] : T(0);
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const int r = wv + 4 * q;
      if (r >= kH) continue;
#pragma unroll
      for (int m = 0; m < kM; ++m) {
        if (lane + 64 * m >= kRow) continue;
        if constexpr (HASW)
          sv[s_lds[m] + r * kW] = pb[q][m] - pw[q][m];
        else
          sv[s_lds[m] + r * kW] = pb[q][m];
      }
    }
  };
  // compute: coarse node (I, J) of the tile, fine node (2 Il + 1, 2 Jl + 1)
  // of the staged plane at its centre
  const int Il = tid % kR2I, Jl = tid / kR2I;
  const int I = tx * kR2I + Il, J = ty * kR2J + Jl;
  const bool own = I < cnnx && J < cnny;
  T acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = T(0);
  const T c2 = x0 ? (T)cf[1] : T(0);
  auto write = [&](int K) {
    if (!own) return;
    const int64_t cn = I + cnnx * (J + (int64_t)cnny * K);
    T* out = bc + (g * cnn + cn) * NV;
#pragma unroll
    for (int v = 0; v < NV; ++v) out[v] = acc[v];
    if (x0) {
      T* xo = x0 + (g * cnn + cn) * NV;
#pragma unroll
      for (int v = 0; v < NV; ++v) xo[v] = c2 * dinv_c[cn * NC + v / KC] * acc[v];
    }
  };
  // planes 2 K0 - 1 .. 2 K1 - 1 that exist
  const int gz_lo = K0 > 0 ? 2 * K0 - 1 : 0, gz_hi = 2 * K1 - 1 < fnz ? 2 * K1 - 1 : fnz;
  fetch(gz_lo);
  for (int gz = gz_lo; gz <= gz_hi; ++gz) {
    stash();
    __syncthreads();
    if (gz < gz_hi) fetch(gz + 1);  // in flight while this plane is summed
    const bool have = gz >= 0 && gz < fnnz;
    T t[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) t[v] = T(0);
    if (have) {
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy) {
        const T wy = T(1) - (T)(dy < 0 ? -dy : dy) / T(2);
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const T wgt = (T(1) - (T)(dx < 0 ? -dx : dx) / T(2)) * wy;
          const int node = (2 * Jl + 1 + dy) * kW + 2 * Il + 1 + dx;
#pragma unroll
          for (int v = 0; v < NV; ++v) t[v] += wgt * sv[v * kH * kW + node];
        }
      }
    }
    __syncthreads();  // the next stash overwrites the plane
    // plane gz = 2K + dz contributes wz t to coarse plane K (dz = -1, 0, 1 in
    // this order, as k_restrict_z sums them)
    if (gz & 1) {
      const int K = (gz - 1) >> 1;  // dz = +1 for K, -1 for K + 1
      if (K >= K0) {
        if (have) {
#pragma unroll
          for (int v = 0; v < NV; ++v) acc[v] += T(0.5) * t[v];
        }
        write(K);
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] = T(0);
      if (have && K + 1 < K1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += T(0.5) * t[v];
      }
    } else {
      const int K = gz >> 1;
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] += T(1) * t[v];
      if (gz + 1 >= fnnz) write(K);  // the last plane: no dz = +1 term
    }
  }
}

// x_f += Z P x_c.  Thread: fine i; row: (j, k, group).  FIX: Dirichlet flags
// of the fine level (no correction on fixed DOFs).  Every load of a row (the
// 8 coarse corners, the fine node, its flags) is unpredicated and issued
// before any is used: a corner outside the interpolation (weight zero) reads
// a clamped node and is dropped by a select, so the sums are those of the
// loop that skips it.  Predicated loads compiled to one branch and one wait
// each.
template <typename T, int NC, int KC, bool FIX>
__global__ void k_prolong_add(int fnx, int fny, int fnz, int fx, int fy, int fz, int groups,
                              const T* __restrict__ xc, const uint8_t* __restrict__ fixed,
                              T* __restrict__ xf) {
  const int fnnx = fnx + 1, fnny = fny + 1, fnnz = fnz + 1;
  const int cnnx = fnx / fx + 1, cnny = fny / fy + 1, cnnz = fnz / fz + 1;
  const int64_t cnn = (int64_t)cnnx * cnny * cnnz, fnn = (int64_t)fnnx * fnny * fnnz;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fnnx) return;
  const int I = i / fx, ax = i - I * fx;
  const T sx = (T)ax / (T)fx;
  const int nrows = fnny * fnnz * groups;
  for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
    const int j = row % fnny, kg = row / fnny, k = kg % fnnz, g = kg / fnnz;
    const int J = j / fy, ay = j - J * fy, K = k / fz, az = k - K * fz;
    const T sy = (T)ay / (T)fy, sz = (T)az / (T)fz;
    const int64_t fnode = i + fnnx * (j + (int64_t)fnny * k);
    const T* xg = xc + g * cnn * NC * KC;
    T* xfn = xf + (g * fnn + fnode) * NC * KC;
    uint8_t fv[NC];
    if constexpr (FIX) {
#pragma unroll
      for (int d = 0; d < NC; ++d) fv[d] = fixed[fnode * NC + d];
    }
    T out[NC * KC];
    load_node<T, NC * KC, KC>(xfn, out);
    T xv[8][NC * KC];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      // the +1 corner exists wherever its weight is nonzero
      const int qx = (q & 1) && ax != 0, qy = ((q >> 1) & 1) && ay != 0, qz = (q >> 2) && az != 0;
      const int64_t cn = (I + qx) + cnnx * ((J + qy) + (int64_t)cnny * (K + qz));
      load_node<T, NC * KC, KC>(xg + cn * NC * KC, xv[q]);
    }
    T acc[NC][KC];
#pragma unroll
    for (int d = 0; d < NC; ++d)
#pragma unroll
      for (int c = 0; c < KC; ++c) acc[d][c] = T(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int qx = q & 1, qy = (q >> 1) & 1, qz = q >> 2;
      const bool skip = (qx && ax == 0) || (qy && ay == 0) || (qz && az == 0);
      const T wgt = (qx ? sx : T(1) - sx) * (qy ? sy : T(1) - sy) * (qz ? sz : T(1) - sz);
#pragma unroll
      for (int d = 0; d < NC; ++d)
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const T a = acc[d][c] + wgt * xv[q][d * KC + c];
          acc[d][c] = skip ? acc[d][c] : a;
        }
    }
#pragma unroll
    for (int d = 0; d < NC; ++d) {
      const bool keep = FIX && fv[d] != 0;
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        const T a = out[d * KC + c] + acc[d][c];
        out[d * KC + c] = keep ? out[d * KC + c] : a;
      }
    }
    store_node<T, NC * KC, KC>(xfn, out);
  }
}

// x = Ainv b on the coarsest level: one wavefront per (row, column group),
// lanes stride the row, fixed-order wave reduction
template <typename T, int KC>
__global__ void __launch_bounds__(256) k_dense_apply(int64_t n, int groups,
                                                     const T* __restrict__ Ainv,
                                                     const T* __restrict__ b,
                                                     T* __restrict__ x) {
  const int lane = threadIdx.x % kWave;
  const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kWave;
  if (wid >= n * groups) return;
  const int64_t row = wid % n, g = wid / n;
  const T* bg = b + g * n * KC;
  double acc[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) acc[c] = 0.0;
  for (int64_t j = lane; j < n; j += kWave) {
    const double a = (double)Ainv[row * n + j];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] += a * (double)bg[j * KC + c];
  }
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    const double v = wave_sum(acc[c]);
    if (lane == 0) x[(g * n + row) * KC + c] = (T)v;
  }
}

// The same product with every column group in one pass over Ainv: b (all G
// groups, G * KC columns) is staged in LDS node-major, and each wavefront
// takes R consecutive rows, so a row of Ainv is read once rather than once
// per group, and each b value read from LDS feeds R rows.  A lane's NJ values
// of its rows (n <= 64 NJ) are loaded into registers before the staging, so
// that every load of the launch is in flight at once.  Per (row, column) the
// lane sums and the wave reduction run in k_dense_apply's order, so the
// results are bit-identical to it.
template <typename T, int KC, int G, int R, int NJ>
__global__ void __launch_bounds__(256) k_dense_apply_rows(int n, const T* __restrict__ Ainv,
                                                          const T* __restrict__ b,
                                                          T* __restrict__ x) {
  constexpr int GK = G * KC;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* bl = reinterpret_cast<T*>(smem_raw);  // [n][GK]
  const int tid = threadIdx.x;
  const int lane = tid % kWave;
  const int row0 = (blockIdx.x * (256 / kWave) + tid / kWave) * R;
  const bool live = row0 < n;
  // unpredicated loads (row and column clamped into the matrix; the clamped
  // values are never used), so that none waits for the one before
  T av[R][NJ];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const T* ar = Ainv + (int64_t)(row0 + r < n ? row0 + r : n - 1) * n;
#pragma unroll
    for (int t = 0; t < NJ; ++t) {
      const int j = lane + kWave * t;
      av[r][t] = ar[j < n ? j : n - 1];
    }
  }
  // b: a thread's NBG values of every group are loaded (again unpredicated,
  // clamped) before any is stored
  constexpr int NBG = (kWave * NJ * KC + 255) / 256;
  const int nk = n * KC;
  T sb[G][NBG];
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int u = 0; u < NBG; ++u) {
      const int i = tid + 256 * u;
      sb[g][u] = b[(int64_t)g * nk + (i < nk ? i : nk - 1)];
    }
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int u = 0; u < NBG; ++u) {
      const int i = tid + 256 * u;
      if (i < nk) {
        const int j = i / KC, c = i - j * KC;
        bl[j * GK + g * KC + c] = sb[g][u];
      }
    }
  __syncthreads();
  if (!live) return;
  double acc[R][GK];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int c = 0; c < GK; ++c) acc[r][c] = 0.0;
#pragma unroll
  for (int t = 0; t < NJ; ++t) {
    const int j = lane + kWave * t;
    if (j < n) {
      double bv[GK];
#pragma unroll
      for (int c = 0; c < GK; ++c) bv[c] = (double)bl[j * GK + c];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double a = (double)av[r][t];
#pragma unroll
        for (int c = 0; c < GK; ++c) acc[r][c] += a * bv[c];
      }
    }
  }
  // wave_sum's butterfly, one step for all R * GK sums at a time, so that
  // their exchanges are in flight together
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1)
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < GK; ++c) acc[r][c] += __shfl_xor(acc[r][c], off, kWave);
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = row0 + r;
      if (row >= n) continue;
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int c = 0; c < KC; ++c) x[((int64_t)g * n + row) * KC + c] = (T)acc[r][g * KC + c];
    }
  }
}

template <typename T, int KC, int R, int NJ>
static bool launch_dense_rows(int n, int groups, const T* Ainv, const T* b, T* x,
                              hipStream_t s) {
  const size_t lds = (size_t)n * groups * KC * sizeof(T);
  const dim3 grid((unsigned)((n + 4 * R - 1) / (4 * R)));
  switch (groups) {
    case 1:
      hipLaunchKernelGGL((k_dense_apply_rows<T, KC, 1, R, NJ>), grid, dim3(256), lds, s, n,
                         Ainv, b, x);
      return true;
    case 2:
      hipLaunchKernelGGL((k_dense_apply_rows<T, KC, 2, R, NJ>), grid, dim3(256), lds, s, n,
                         Ainv, b, x);
      return true;
    case 4:
      hipLaunchKernelGGL((k_dense_apply_rows<T, KC, 4, R, NJ>), grid, dim3(256), lds, s, n,
                         Ainv, b, x);
      return true;
    default:
      return false;
  }
}

// SKT_MG_DENSE_APPLY=wave: one wavefront per (row, group) at every size
static bool dense_rows_enabled() {
  const char* v = std::getenv("SKT_MG_DENSE_APPLY");
  return !(v && std::strcmp(v, "wave") == 0);
}

// x = Ainv b for `groups` column groups of KC: the row-blocked kernel when b
// fits in 64 KB of LDS, n <= 2048 and the group count is 1, 2 or 4, else one
// wavefront per (row, group)
template <typename T, int KC>
static void launch_dense_apply(int64_t n, int groups, const T* Ainv, const T* b, T* x,
                               hipStream_t s) {
  const size_t lds = (size_t)n * groups * KC * sizeof(T);
  if (lds <= 65536 && n <= 64 * 32 && dense_rows_enabled()) {
    if (n <= 64 * 8 ? launch_dense_rows<T, KC, 2, 8>((int)n, groups, Ainv, b, x, s)
                    : launch_dense_rows<T, KC, 2, 32>((int)n, groups, Ainv, b, x, s))
      return;
  }
  const int64_t waves = n * groups;
  hipLaunchKernelGGL((k_dense_apply<T, KC>), dim3((unsigned)((waves * kWave + 255) / 256)),
                     dim3(256), 0, s, n, groups, Ainv, b, x);
}

// ------------------------------------------------- smoother bounds (power)

__device__ __forceinline__ double hash_unit(int64_t i) {
  uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

template <typename T>
__global__ void k_pow_init(int64_t n, const T* __restrict__ dinv, T* __restrict__ v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = dinv[i] != T(0) ? (T)hash_unit(i) : T(0);
}

// y <- dinv y (if scale_by_dinv), partial sums of y^2 per block
template <typename T>
__global__ void k_pow_norm(int64_t n, const T* __restrict__ dinv, T* __restrict__ y,
                           int scale_by_dinv, double* __restrict__ part) {
  __shared__ double sh[kBlock / kWave];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T v = y[i];
    if (scale_by_dinv) { v *= dinv[i]; y[i] = v; }
    acc += (double)v * (double)v;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// norm = sqrt(sum part); inv_out = 1 / norm.  With lam_out, lam = norm *
// inv_in (inv_in: 1 / the norm of the vector the operator was applied to).
__global__ void k_pow_finish(int nb, const double* __restrict__ part, double* __restrict__ lam_out,
                             const double* __restrict__ inv_in, double* __restrict__ inv_out) {
  __shared__ double sh[kBlock / kWave];
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) a += part[i];
  const double s = block_sum(a, sh);
  if (threadIdx.x == 0) {
    const double nrm = sqrt(s);
    if (lam_out) *lam_out = nrm * *inv_in;
    *inv_out = nrm > 0.0 ? 1.0 / nrm : 0.0;
  }
}

// v = dinv y (an unnormalised power step)
template <typename T>
__global__ void k_pow_dscale(int64_t n, const T* __restrict__ dinv, const T* __restrict__ y,
                             T* __restrict__ v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = dinv[i] * y[i];
}

template <typename T>
__global__ void k_pow_scale(int64_t n, const double* __restrict__ inv, const T* __restrict__ y,
                            T* __restrict__ v) {
  const T s = (T)*inv;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = y[i] * s;
}

// Chebyshev (c1, c2) per step for the interval [beta/ratio, beta],
// beta = 1.1 * lambda estimate
__global__ void k_cheb_coef(int nlev, int degree, double ratio, const double* __restrict__ lam,
                            double* __restrict__ coef) {
  const int l = threadIdx.x;
  if (l >= nlev) return;
  const double bmax = 1.1 * lam[l];
  const double amin = bmax / ratio;
  const double theta = 0.5 * (bmax + amin), delta = 0.5 * (bmax - amin);
  const double sigma = theta / delta;
  double* c = coef + (size_t)l * 2 * degree;
  double rho_old = 1.0 / sigma;
  c[0] = 0.0;
  c[1] = 1.0 / theta;
  for (int s = 1; s < degree; ++s) {
    const double rho = 1.0 / (2.0 * sigma - rho_old);
    c[2 * s] = rho * rho_old;
    c[2 * s + 1] = 2.0 * rho / delta;
    rho_old = rho;
  }
}

// ------------------------------------------------------------ host side

static size_t tsize(int dtype) { return dtype == SKT_F64 ? sizeof(double) : sizeof(float); }

void mg_free(MGData* mg) {
  // the set-up's side work and any background inverse are done
  if (mg->side) (void)hipStreamSynchronize(mg->side);
  if (mg->side2) (void)hipStreamSynchronize(mg->side2);
  if (mg->gjs) {
    (void)hipStreamSynchronize(mg->gjs);
    (void)hipStreamDestroy(mg->gjs);
    (void)hipEventDestroy(mg->ev_asm);
    (void)hipEventDestroy(mg->ev_gj);
    mg->gjs = nullptr;
    mg->ev_asm = mg->ev_gj = nullptr;
  }
  mg->have_inv = 0;
  mg->inv_pending = 0;
  for (hipEvent_t& e : mg->tev)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  mg->tcount = 0;
  mg->time_fine = 0;
  for (MGLevel& l : mg->lv) {
    (void)hipFree(l.Kel); (void)hipFree(l.st); (void)hipFree(l.dinv);
    (void)hipFree(l.b); (void)hipFree(l.x); (void)hipFree(l.w); (void)hipFree(l.d);
    (void)hipFree(l.rt);
    l.Kel = nullptr; l.st = l.dinv = l.b = l.x = l.w = l.d = l.rt = nullptr;
  }
  (void)hipFree(mg->Ad); (void)hipFree(mg->Ad2); (void)hipFree(mg->Ainv);
  (void)hipFree(mg->lam); (void)hipFree(mg->coef); (void)hipFree(mg->part); (void)hipFree(mg->pv);
  (void)hipFree(mg->pvw);
  (void)hipFree(mg->K0b);
  if (mg->side) {
    (void)hipStreamDestroy(mg->side);
    (void)hipStreamDestroy(mg->side2);
    (void)hipEventDestroy(mg->ev_fork);
    (void)hipEventDestroy(mg->ev_gal);
    (void)hipEventDestroy(mg->ev_st);
    (void)hipEventDestroy(mg->ev_join);
    mg->side = mg->side2 = nullptr;
    mg->ev_fork = mg->ev_gal = mg->ev_st = mg->ev_join = nullptr;
  }
  (void)hipFree(mg->Mch);
  (void)hipFree(mg->gflag);
  mg->Mch = nullptr;
  mg->gflag = nullptr;
  mg->Ad = mg->Ad2 = nullptr; mg->Ainv = nullptr; mg->lam = mg->coef = mg->part = nullptr;
  mg->K0b = nullptr;
  mg->pv = nullptr;
  mg->pvw = nullptr;
  mg->warm = 0;
  mg->lv.clear();
  mg->dtype = -1;
}

// Stencil layout of a level with nn nodes: big tiles from kTiledNodes
// nodes, small tiles below.  SKT_MG_STENCIL_LAYOUT=node|big|small forces one
// layout on every level (read when a hierarchy is allocated; the three sum
// in the same order, which a test checks).
static int stencil_layout(int64_t nn) {
  const char* v = std::getenv("SKT_MG_STENCIL_LAYOUT");
  if (v && std::strcmp(v, "node") == 0) return ST_NODE;
  if (v && std::strcmp(v, "big") == 0) return ST_TILE_BIG;
  if (v && std::strcmp(v, "small") == 0) return ST_TILE_SMALL;
  return nn >= kTiledNodes ? ST_TILE_BIG : ST_TILE_SMALL;
}

// Power-iteration steps from the previous set-up's vector;
// SKT_MG_POWER_WARM=n (1..64) overrides the default.
static int warm_power_steps(const MGData* mg) {
  const char* v = std::getenv("SKT_MG_POWER_WARM");
  if (v) {
    const int n = std::atoi(v);
    if (n >= 1 && n <= 64) return n;
  }
  return mg->power_iters_warm;
}

int mg_warm_steps(const MGData* mg) { return warm_power_steps(mg); }

// SKT_MG_CHEB0=0 runs a degree-2 fine-level smoothing as the separate first
// step plus one fused step instead of the single two-step pass (the same
// arithmetic; a test checks that the solves agree bit for bit).
static bool cheb0_enabled() {
  const char* v = std::getenv("SKT_MG_CHEB0");
  return !(v && v[0] == '0' && v[1] == 0);
}

// SKT_MG_COARSE_REFRESH=n (1..64): set-ups one coarsest inverse may serve.
int mg_coarse_refresh(const MGData* mg) {
  const char* v = std::getenv("SKT_MG_COARSE_REFRESH");
  if (v) {
    const int n = std::atoi(v);
    if (n >= 1 && n <= 64) return n;
  }
  return mg->coarse_refresh;
}

static int lag_enabled() {
  const char* v = std::getenv("SKT_MG_LAG");
  return (v && v[0] == '1' && v[1] == 0) ? 1 : 0;
}

int mg_alloc(MGData* mg, const skt_operator* op, const MGConfig& cfg, int dtype, int max_cols) {
  if (mg->dtype == dtype && mg->max_cols == max_cols) return SKT_OK;
  SKT_CHECK((op->nx + 1) * (op->ny + 1) < INT32_MAX &&
                (op->ny + 1) * (op->nz + 1) * (int64_t)max_cols < INT32_MAX,
            "grid too large for the multigrid transfers' 32-bit row indices");
  mg_free(mg);
  std::vector<MGLevel> lv;
  if (!mg_plan(op, cfg, &lv)) {
    set_error("grid cannot be coarsened for multigrid");
    return SKT_EINVAL;
  }
  const int NC = op->ncomp, ND = 8 * NC;
  const size_t ts = tsize(dtype);
  const int L = (int)lv.size() - 1;
  for (int l = 0; l <= L; ++l) {
    MGLevel& v = lv[l];
    const size_t nv = (size_t)v.ndof * max_cols;
    if (l == 0) {
      SKT_HIP(hipMalloc(&v.x, ts * nv));
      continue;
    }
    SKT_HIP(hipMalloc(&v.Kel, sizeof(double) * (size_t)(v.nx * v.ny * v.nz) * ND * ND));
    SKT_HIP(hipMalloc(&v.rt, ts * (size_t)(v.nx + 1) * (v.ny + 1) * (lv[l - 1].nz + 1) * NC *
                                 max_cols));
    SKT_HIP(hipMalloc(&v.b, ts * nv));
    SKT_HIP(hipMalloc(&v.x, ts * nv));
    if (l < L) {
      v.layout = stencil_layout(v.nn);
      SKT_HIP(hipMalloc(&v.st, ts * (size_t)st_nodes(v.nx, v.ny, v.nz, v.layout) * 27 * NC * NC));
      SKT_HIP(hipMalloc(&v.dinv, ts * v.ndof));
      SKT_HIP(hipMalloc(&v.w, ts * nv));
      SKT_HIP(hipMalloc(&v.d, ts * nv));
    }
  }
  const int64_t nc = lv[L].ndof;
  SKT_HIP(hipMalloc(&mg->Ad, sizeof(double) * nc * nc));
  SKT_HIP(hipMalloc(&mg->Ad2, sizeof(double) * 2 * kGJ * kGJ));
  SKT_HIP(hipMalloc(&mg->Ainv, ts * nc * nc));
  SKT_HIP(hipMalloc(&mg->lam, sizeof(double) * (L + 1) * 2));
  SKT_HIP(hipMalloc(&mg->coef, sizeof(double) * (L + 1) * 2 * std::max(cfg.degree, 1)));
  SKT_HIP(hipMalloc(&mg->part, sizeof(double) * 1024));
  SKT_HIP(hipMalloc(&mg->pv, ts * 2 * lv[0].ndof));
  int64_t npw = 0;
  for (int l = 0; l < L; ++l) npw += lv[l].ndof;
  SKT_HIP(hipMalloc(&mg->pvw, ts * npw));
  mg->warm = 0;
  // unit element matrix in corner-bit order
  std::vector<double> K0b((size_t)ND * ND);
  for (int A = 0; A < 8; ++A)
    for (int B = 0; B < 8; ++B) {
      const int a = corner_id_h(A & 1, (A >> 1) & 1, A >> 2);
      const int b = corner_id_h(B & 1, (B >> 1) & 1, B >> 2);
      for (int d = 0; d < NC; ++d)
        for (int e = 0; e < NC; ++e)
          K0b[(A * NC + d) * ND + B * NC + e] = op->K0h[(a * NC + d) * ND + b * NC + e];
    }
  SKT_HIP(hipMalloc(&mg->K0b, sizeof(double) * K0b.size()));
  SKT_HIP(hipMemcpy(mg->K0b, K0b.data(), sizeof(double) * K0b.size(), hipMemcpyHostToDevice));
  mg->lv = lv;
  mg->dtype = dtype;
  ++mg->gen;
  mg->max_cols = max_cols;
  mg->degree = std::max(cfg.degree, 1);
  mg->ratio = cfg.ratio;
  if (const char* rv = std::getenv("SKT_MG_RATIO")) {  // smoothing interval [lambda/r, lambda]
    const double r = std::atof(rv);
    if (r > 1.0 && r < 1000.0) mg->ratio = r;
  }
  mg->power_iters = cfg.power_iters;
  mg->lag = lag_enabled();
  return SKT_OK;
}

// apply of level l (k columns, group kc) in T
template <typename T, int MODE = ST_APPLY>
static int level_apply(MGData* mg, const skt_operator* op, const void* E, int l, int k, int kc,
                       const T* x, T* y, hipStream_t s, StOut<T> oa = StOut<T>()) {
  if (l == 0) return op_apply(const_cast<skt_operator*>(op), mg->dtype, k, kc, E, x, y, s);
  const MGLevel& v = mg->lv[l];
  const int groups = k / kc;
  const T* st = (const T*)v.st;
  // tile-major levels: LDS-tiled kernel of the level's tile shape (x staged
  // once per tile); node-major: one (node, row, column group) per thread
  constexpr int GMT = sizeof(T) == 4 ? 4 : 2;  // <= 62 KB of LDS for NC * KC = 6
  const unsigned gmb = (unsigned)((groups + GMT - 1) / GMT);
  const unsigned g = grid_for(v.nn * op->ncomp * groups);
#define SKT_SA(NC_, KC_)                                                                         \
  do {                                                                                           \
    if (v.layout == ST_TILE_BIG) {                                                               \
      const unsigned gt = (unsigned)(st_tiles(v.nx, v.ny, v.nz, kTileBig) * gmb);                \
      hipLaunchKernelGGL((k_stencil_apply_tiled<T, NC_, KC_, GMT, MODE, kTileBig.x, kTileBig.y,   \
                                                kTileBig.z>),                                    \
                         dim3(gt), dim3(kTileBig.x * kTileBig.y * kTileBig.z * NC_), 0, s, v.nx,  \
                         v.ny, v.nz, groups, st, x, y, oa);                                      \
    } else if (v.layout == ST_TILE_SMALL) {                                                      \
      const unsigned gt = (unsigned)(st_tiles(v.nx, v.ny, v.nz, kTileSmall) * gmb);              \
      hipLaunchKernelGGL((k_stencil_apply_tiled<T, NC_, KC_, GMT, MODE, kTileSmall.x,             \
                                                kTileSmall.y, kTileSmall.z>),                    \
                         dim3(gt), dim3(kTileSmall.x * kTileSmall.y * kTileSmall.z * NC_), 0, s,  \
                         v.nx, v.ny, v.nz, groups, st, x, y, oa);                                \
    } else {                                                                                     \
      hipLaunchKernelGGL((k_stencil_apply<T, NC_, KC_, 1, MODE>), dim3(g), dim3(kBlock), 0, s,    \
                         v.nx, v.ny, v.nz, groups, st, x, y, oa);                                \
    }                                                                                            \
  } while (0)
  if (op->ncomp == 3) {
    if (kc == 1) SKT_SA(3, 1);
    else if (kc == 2) SKT_SA(3, 2);
    else { set_error("unsupported column block"); return SKT_EINVAL; }
  } else {
    if (kc == 1) SKT_SA(1, 1);
    else if (kc == 2) SKT_SA(1, 2);
    else if (kc == 4) SKT_SA(1, 4);
    else { set_error("unsupported column block"); return SKT_EINVAL; }
  }
#undef SKT_SA
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

// One fused Chebyshev-Jacobi step on a stencil level: x_out = x + c1 (x -
// x_out_before) + c2 dinv (b - S x).
template <typename T>
static int level_cheb(MGData* mg, const skt_operator* op, int l, int k, int kc, const T* x,
                      T* xout, const double* coef, int prev_zero, hipStream_t s,
                      const double* stop) {
  const MGLevel& v = mg->lv[l];
  StOut<T> oa;
  oa.b = (const T*)v.b;
  oa.dinv = (const T*)v.dinv;
  oa.coef = coef;
  oa.prev_zero = prev_zero;
  oa.stop = stop;
  return level_apply<T, ST_CHEB>(mg, op, nullptr, l, k, kc, x, xout, s, oa);
}

// In-place blocked Gauss-Jordan inverse of the coarsest matrix in Ad, on
// stream g (pivot block inverses alternate between the two halves of Ad2).
static void queue_gj(MGData* mg, hipStream_t g) {
  const int64_t n = mg->lv.back().ndof;
  const unsigned tiles = (unsigned)((n + kTile - 1) / kTile);
  const unsigned strips = (unsigned)((n + kTileRC - 1) / kTileRC);
  hipLaunchKernelGGL(k_gj_block_inv, dim3(1), dim3(256), 0, g, n, (int64_t)0,
                     (int)std::min<int64_t>(kGJ, n), mg->Ad, mg->Ad2);
  int par = 0;
  for (int64_t b0 = 0; b0 < n; b0 += kGJ, par ^= 1) {
    const int nb = (int)std::min<int64_t>(kGJ, n - b0);
    const int64_t b1 = b0 + kGJ;
    const int nb1 = b1 < n ? (int)std::min<int64_t>(kGJ, n - b1) : 0;
    double* bi = mg->Ad2 + par * kGJ * kGJ;
    double* bi1 = mg->Ad2 + (par ^ 1) * kGJ * kGJ;
    hipLaunchKernelGGL(k_gj_rows, dim3(strips), dim3(256), 0, g, n, b0, nb, bi, mg->Ad);
    hipLaunchKernelGGL(k_gj_update, dim3(tiles, tiles), dim3(256), 0, g, n, b0, nb, mg->Ad);
    hipLaunchKernelGGL(k_gj_cols_next, dim3(strips + 1), dim3(256), 0, g, n, b0, nb,
                       (const double*)bi, b1, nb1, bi1, mg->Ad);
  }
}

bool mg_lagged(const MGData* mg) { return mg->lag && mg->have_inv; }

int mg_set_timing(MGData* mg, int enable) {
  if (enable && !mg->tev[0])
    for (hipEvent_t& e : mg->tev) SKT_HIP(hipEventCreate(&e));
  mg->time_fine = enable ? 1 : 0;
  return SKT_OK;
}

int mg_collect_timing(MGData* mg) {
  for (int i = 0; i < mg->tcount; ++i) {
    if (mg->ttag[i] < 0) continue;
    float ms = 0.f;
    SKT_HIP(hipEventElapsedTime(&ms, mg->tev[2 * i], mg->tev[2 * i + 1]));
    mg->tms[mg->ttag[i]] += ms;
    mg->tn[mg->ttag[i]] += 1;
  }
  mg->tcount = 0;
  return SKT_OK;
}

// event pair around one fine-level launch when timing (at most kTimedMax
// between two collections; later launches go untimed)
static int fine_timer_begin(MGData* mg, int tag, hipStream_t) {
  if (!mg->time_fine || mg->tcount >= MGData::kTimedMax) return -1;
  const int i = mg->tcount++;
  mg->ttag[i] = tag;
  modal_time_next(mg->tev[2 * i], mg->tev[2 * i + 1]);
  return i;
}

// a request that no launch took leaves its events unrecorded: drop the slot
static void fine_timer_end(MGData* mg, int i, hipStream_t) {
  if (i >= 0 && modal_time_clear()) mg->ttag[i] = -1;
}

int mg_lag_pre(MGData* mg, hipStream_t s) {
  if (!mg->inv_pending) return SKT_OK;
  const int64_t n = mg->lv.back().ndof;
  SKT_HIP(hipStreamWaitEvent(s, mg->ev_gj, 0));
  if (mg->dtype == SKT_F64)
    hipLaunchKernelGGL(k_cast<double>, dim3(grid_for(n * n)), dim3(kBlock), 0, s, n * n,
                       (const double*)mg->Ad, (double*)mg->Ainv);
  else
    hipLaunchKernelGGL(k_cast<float>, dim3(grid_for(n * n)), dim3(kBlock), 0, s, n * n,
                       (const double*)mg->Ad, (float*)mg->Ainv);
  SKT_LAUNCH_CHECK();
  mg->inv_pending = 0;
  return SKT_OK;
}

int mg_lag_post(MGData* mg, hipStream_t s) {
  if (!mg->gjs) {
    SKT_HIP(hipStreamCreateWithFlags(&mg->gjs, hipStreamNonBlocking));
    SKT_HIP(hipEventCreateWithFlags(&mg->ev_asm, hipEventDisableTiming));
    SKT_HIP(hipEventCreateWithFlags(&mg->ev_gj, hipEventDisableTiming));
  }
  SKT_HIP(hipEventRecord(mg->ev_asm, s));
  SKT_HIP(hipStreamWaitEvent(mg->gjs, mg->ev_asm, 0));
  queue_gj(mg, mg->gjs);
  SKT_LAUNCH_CHECK();
  SKT_HIP(hipEventRecord(mg->ev_gj, mg->gjs));
  mg->inv_pending = 1;
  return SKT_OK;
}

template <typename T>
static int setup_t(MGData* mg, const skt_operator* op, const double* E64, const void* E,
                   const void* dinv0, hipStream_t s) {
  const int NC = op->ncomp, ND = 8 * NC;
  const int L = (int)mg->lv.size() - 1;
  const double* dK0 = mg->K0b;
  // The hierarchy is built on side streams while the main stream runs the
  // power iterations: `side` takes the Galerkin products and then the
  // coarsest inverse, `side2` the stencils and Jacobi inverses once the
  // products they need are done (ev_gal).  The fine level's power iteration
  // needs only the fine operator, the coarse levels' wait for the stencils
  // (ev_st), and the set-up joins the coarsest inverse (ev_join) before
  // returning.  The side streams touch the level matrices, stencils, Jacobi
  // inverses and the coarsest buffers only.
  if (!mg->side) {
    SKT_HIP(hipStreamCreateWithFlags(&mg->side, hipStreamNonBlocking));
    SKT_HIP(hipStreamCreateWithFlags(&mg->side2, hipStreamNonBlocking));
    SKT_HIP(hipEventCreateWithFlags(&mg->ev_fork, hipEventDisableTiming));
    SKT_HIP(hipEventCreateWithFlags(&mg->ev_gal, hipEventDisableTiming));
    SKT_HIP(hipEventCreateWithFlags(&mg->ev_st, hipEventDisableTiming));
    SKT_HIP(hipEventCreateWithFlags(&mg->ev_join, hipEventDisableTiming));
  }
  SKT_HIP(hipEventRecord(mg->ev_fork, s));
  SKT_HIP(hipStreamWaitEvent(mg->side, mg->ev_fork, 0));
  hipStream_t g = mg->side;
  // Galerkin hierarchy.  The first coarse level takes E_j M_j for its
  // interior elements (M_j and the boundary flags depend on the operator
  // only, so they are built once) and the masked product for the rest.
  if (op->mesh == SKT_MESH_TET4) {
    const MGLevel& c = mg->lv[1];
    const int64_t nce = c.nx * c.ny * c.nz;
    const uint8_t* fx0 = op->n_fixed > 0 ? op->fixed : nullptr;
    if (NC == 3)
      hipLaunchKernelGGL((k_galerkin_tet<3>), dim3((unsigned)nce), dim3(64), 0, g, c.nx, c.ny,
                         c.nz, c.f[0], c.f[1], c.f[2], op->conn, op->g64, op->vol64, op->lam0,
                         op->mu0, op->n_el, E64, fx0, c.Kel);
    else
      hipLaunchKernelGGL((k_galerkin_tet<1>), dim3((unsigned)nce), dim3(64), 0, g, c.nx, c.ny,
                         c.nz, c.f[0], c.f[1], c.f[2], op->conn, op->g64, op->vol64, op->lam0,
                         op->mu0, op->n_el, E64, fx0, c.Kel);
    SKT_LAUNCH_CHECK();
  } else {
    const MGLevel& c = mg->lv[1];
    const int nch = c.f[0] * c.f[1] * c.f[2];
    const int64_t nce = c.nx * c.ny * c.nz;
    if (!mg->Mch) {
      SKT_HIP(hipMalloc(&mg->Mch, sizeof(double) * (size_t)nch * ND * ND));
      if (NC == 3)
        hipLaunchKernelGGL((k_galerkin_children<3>), dim3(nch), dim3(64), 0, g, c.f[0], c.f[1],
                           c.f[2], dK0, mg->Mch);
      else
        hipLaunchKernelGGL((k_galerkin_children<1>), dim3(nch), dim3(64), 0, g, c.f[0], c.f[1],
                           c.f[2], dK0, mg->Mch);
      if (op->n_fixed > 0) {
        SKT_HIP(hipMalloc(&mg->gflag, (size_t)nce));
        if (NC == 3)
          hipLaunchKernelGGL((k_galerkin_flags<3>), dim3(grid_for(nce)), dim3(kBlock), 0, g,
                             c.nx, c.ny, c.nz, c.f[0], c.f[1], c.f[2], op->fixed, mg->gflag);
        else
          hipLaunchKernelGGL((k_galerkin_flags<1>), dim3(grid_for(nce)), dim3(kBlock), 0, g,
                             c.nx, c.ny, c.nz, c.f[0], c.f[1], c.f[2], op->fixed, mg->gflag);
      }
      SKT_LAUNCH_CHECK();
    }
    SKT_CHECK(nch <= kMaxChildren, "internal: coarsening factor above 5");
    const unsigned gi = (unsigned)std::min<int64_t>(nce, 8192);
    if (NC == 3)
      hipLaunchKernelGGL((k_galerkin_interior<3>), dim3(gi), dim3(64 * 9), 0, g, c.nx, c.ny, c.nz,
                         c.f[0], c.f[1], c.f[2], mg->Mch, E64, mg->gflag, c.Kel);
    else
      hipLaunchKernelGGL((k_galerkin_interior<1>), dim3(gi), dim3(64), 0, g, c.nx, c.ny, c.nz,
                         c.f[0], c.f[1], c.f[2], mg->Mch, E64, mg->gflag, c.Kel);
    if (mg->gflag) {
      if (NC == 3)
        hipLaunchKernelGGL((k_galerkin<3>), dim3((unsigned)nce), dim3(64), 0, g, c.nx, c.ny,
                           c.nz, c.f[0], c.f[1], c.f[2], nullptr, dK0, E64, op->fixed,
                           mg->gflag, c.Kel);
      else
        hipLaunchKernelGGL((k_galerkin<1>), dim3((unsigned)nce), dim3(64), 0, g, c.nx, c.ny,
                           c.nz, c.f[0], c.f[1], c.f[2], nullptr, dK0, E64, op->fixed,
                           mg->gflag, c.Kel);
    }
    SKT_LAUNCH_CHECK();
  }
  for (int l = 2; l <= L; ++l) {
    // the stencil levels' products are done: side2 may start
    if (l == L) SKT_HIP(hipEventRecord(mg->ev_gal, g));
    const MGLevel& c = mg->lv[l];
    const int64_t nce = c.nx * c.ny * c.nz;
    const double* Kf = mg->lv[l - 1].Kel;
    if (NC == 3)
      hipLaunchKernelGGL((k_galerkin<3>), dim3((unsigned)nce), dim3(64), 0, g, c.nx, c.ny, c.nz,
                         c.f[0], c.f[1], c.f[2], Kf, dK0, E64, nullptr, nullptr, c.Kel);
    else
      hipLaunchKernelGGL((k_galerkin<1>), dim3((unsigned)nce), dim3(64), 0, g, c.nx, c.ny, c.nz,
                         c.f[0], c.f[1], c.f[2], Kf, dK0, E64, nullptr, nullptr, c.Kel);
    SKT_LAUNCH_CHECK();
  }
  if (L < 2) SKT_HIP(hipEventRecord(mg->ev_gal, g));
  // stencils and Jacobi inverses of the intermediate levels (side2)
  SKT_HIP(hipStreamWaitEvent(mg->side2, mg->ev_gal, 0));
  g = mg->side2;
  for (int l = 1; l < L; ++l) {
    const MGLevel& v = mg->lv[l];
    const int64_t tot = st_nodes(v.nx, v.ny, v.nz, v.layout) * 27 * NC * NC;
    if (NC == 3) {
      hipLaunchKernelGGL((k_stencil_asm<T, 3>), dim3(grid_for(tot)), dim3(kBlock), 0, g, v.nx,
                         v.ny, v.nz, v.layout, v.Kel, (T*)v.st);
      hipLaunchKernelGGL((k_stencil_dinv<T, 3>), dim3(grid_for(v.ndof)), dim3(kBlock), 0, g,
                         v.nx, v.ny, v.nz, v.layout, (const T*)v.st, (T*)v.dinv);
    } else {
      hipLaunchKernelGGL((k_stencil_asm<T, 1>), dim3(grid_for(tot)), dim3(kBlock), 0, g, v.nx,
                         v.ny, v.nz, v.layout, v.Kel, (T*)v.st);
      hipLaunchKernelGGL((k_stencil_dinv<T, 1>), dim3(grid_for(v.ndof)), dim3(kBlock), 0, g,
                         v.nx, v.ny, v.nz, v.layout, (const T*)v.st, (T*)v.dinv);
    }
    SKT_LAUNCH_CHECK();
  }
  SKT_HIP(hipEventRecord(mg->ev_st, g));
  g = mg->side;
  // coarsest: dense matrix, blocked Gauss-Jordan inverse, cast to T (side
  // stream).  Enqueued after the fine level's power iteration, so that the
  // host's many small launches for it do not hold back the main stream.
  // With a lagged inverse (mg_lagged) only the matrix is assembled here;
  // mg_lag_post inverts it in the background after the set-up.
  const bool lagged = mg_lagged(mg);
  const bool reuse = mg->reuse_inv && mg->have_inv && !lagged;
  bool coarsest_queued = false;
  auto queue_coarsest = [&]() -> int {
    coarsest_queued = true;
    const MGLevel& c = mg->lv[L];
    const int64_t n = c.ndof;
    if (reuse) {  // the previous inverse stays in Ainv
      SKT_HIP(hipEventRecord(mg->ev_join, g));
      return SKT_OK;
    }
    if (NC == 3)
      hipLaunchKernelGGL((k_dense_asm<3>), dim3(grid_for(n * n)), dim3(kBlock), 0, g, c.nx, c.ny,
                         c.nz, c.Kel, mg->Ad);
    else
      hipLaunchKernelGGL((k_dense_asm<1>), dim3(grid_for(n * n)), dim3(kBlock), 0, g, c.nx, c.ny,
                         c.nz, c.Kel, mg->Ad);
    hipLaunchKernelGGL(k_dense_fix_diag, dim3(grid_for(n)), dim3(kBlock), 0, g, n, mg->Ad);
    if (!lagged) {
      queue_gj(mg, g);
      hipLaunchKernelGGL(k_cast<T>, dim3(grid_for(n * n)), dim3(kBlock), 0, g, n * n, mg->Ad,
                         (T*)mg->Ainv);
    }
    SKT_LAUNCH_CHECK();
    SKT_HIP(hipEventRecord(mg->ev_join, g));
    return SKT_OK;
  };
  // smoother bounds: power iteration on D^-1 A (one column) per level
  double* inv = mg->lam + (L + 1);  // scratch scalar slot
  int64_t off = 0;
  bool st_joined = false;
  for (int l = 0; l < L; ++l) {
    if (l == 1) {
      if (int stt = queue_coarsest()) return stt;
      SKT_HIP(hipStreamWaitEvent(s, mg->ev_st, 0));
      st_joined = true;
    }
    const MGLevel& v = mg->lv[l];
    const int64_t n = v.ndof;
    T* pv0 = (T*)mg->pvw + off;      // this level's vector, kept for the next set-up
    T* pv1 = (T*)mg->pv;
    off += n;
    const T* dinv = l == 0 ? (const T*)dinv0 : (const T*)v.dinv;
    const unsigned nb = std::min<unsigned>(grid_for(n), 1024u);
    if (!mg->warm)
      hipLaunchKernelGGL(k_pow_init<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, dinv, pv0);
    // The iterate is left unnormalised between steps (it grows by about
    // lambda per step, far from the float range over these few steps); the
    // last step's input and output norms give lambda, and the normalised
    // result is kept for the next set-up.
    const int steps = mg->warm ? warm_power_steps(mg) : mg->power_iters;
    for (int it = 0; it < steps; ++it) {
      const bool last = it + 1 == steps;
      if (last) {
        hipLaunchKernelGGL(k_pow_norm<T>, dim3(nb), dim3(kBlock), 0, s, n, dinv, pv0, 0,
                           mg->part);
        hipLaunchKernelGGL(k_pow_finish, dim3(1), dim3(kBlock), 0, s, (int)nb, mg->part,
                           (double*)nullptr, (const double*)nullptr, inv);
      }
      int stt = level_apply<T>(mg, op, E, l, 1, 1, pv0, pv1, s);
      if (stt) return stt;
      if (!last) {
        hipLaunchKernelGGL(k_pow_dscale<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, dinv, pv1,
                           pv0);
        continue;
      }
      hipLaunchKernelGGL(k_pow_norm<T>, dim3(nb), dim3(kBlock), 0, s, n, dinv, pv1, 1, mg->part);
      hipLaunchKernelGGL(k_pow_finish, dim3(1), dim3(kBlock), 0, s, (int)nb, mg->part,
                         mg->lam + l, (const double*)inv, inv);
      hipLaunchKernelGGL(k_pow_scale<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, inv, pv1,
                         pv0);
    }
    SKT_LAUNCH_CHECK();
  }
  if (!coarsest_queued) {
    if (int stt = queue_coarsest()) return stt;
  }
  mg->warm = 1;
  hipLaunchKernelGGL(k_cheb_coef, dim3(1), dim3(64), 0, s, L, mg->degree, mg->ratio, mg->lam,
                     mg->coef);
  SKT_LAUNCH_CHECK();
  // every side-stream branch joins the main stream before the set-up
  // returns (a graph capture of the set-up requires it)
  if (!st_joined) SKT_HIP(hipStreamWaitEvent(s, mg->ev_st, 0));
  SKT_HIP(hipStreamWaitEvent(s, mg->ev_join, 0));
  if (!lagged) mg->have_inv = 1;
  return SKT_OK;
}

int mg_setup(MGData* mg, const skt_operator* op, const double* E64, const void* E,
             const void* dinv0, hipStream_t s) {
  SKT_CHECK(mg->dtype >= 0 && !mg->lv.empty(), "multigrid not allocated");
  if (mg->dtype == SKT_F64) return setup_t<double>(mg, op, E64, E, dinv0, s);
  return setup_t<float>(mg, op, E64, E, dinv0, s);
}

template <typename T, int NC, int KC>
static void launch_restrict(const MGLevel& f, const MGLevel& c, int groups, const T* b,
                            const T* w, const uint8_t* fixed, T* bc, hipStream_t s) {
  T* tbuf = (T*)c.rt;
  const RowGrid gxy = row_grid(c.nx + 1, (c.ny + 1) * (f.nz + 1) * groups);
  hipLaunchKernelGGL((k_restrict_xy<T, NC, KC>), gxy.grid, gxy.block, 0, s, (int)f.nx, (int)f.ny,
                     (int)f.nz, c.f[0], c.f[1], groups, b, w, fixed, tbuf);
  const RowGrid gz = row_grid((c.nx + 1) * (c.ny + 1), (c.nz + 1) * groups);
  hipLaunchKernelGGL((k_restrict_z<T, NC, KC>), gz.grid, gz.block, 0, s, (int)(c.nx + 1),
                     (int)(c.ny + 1), (int)f.nz, c.f[2], groups, (const T*)tbuf, bc);
}

// SKT_MG_TRANSFERS=rowgrid forces the row-grid restriction on every level
// (the factor-2 kernel gives the same sums; a test checks that the solves
// agree bit for bit).
static bool fused_transfers_enabled() {
  const char* v = std::getenv("SKT_MG_TRANSFERS");
  return !(v && std::strcmp(v, "rowgrid") == 0);
}

// Fused restriction into level c when every factor is 2 (k_restrict222; b
// zero at Dirichlet DOFs); with x0, also the first smoothing iterate of c
// (coefficients cf).  Returns false, launching nothing, otherwise.
template <typename T, int NC, int KC>
static bool launch_restrict_fused(const MGLevel& f, const MGLevel& c, int groups, const T* b,
                                  const T* w, T* bc, const double* cf, T* x0, hipStream_t s) {
  if (c.f[0] != 2 || c.f[1] != 2 || c.f[2] != 2 || !fused_transfers_enabled()) return false;
  const int tiles_x = (int)((c.nx + 1 + kR2I - 1) / kR2I);
  const int tiles_y = (int)((c.ny + 1 + kR2J - 1) / kR2J);
  const int cnnz = (int)(c.nz + 1);
  // planes per chunk: halved until there are >= 1024 workgroups (each chunk
  // streams its 2 zlen + 1 fine planes one after the other)
  int zlen = 4;
  const int64_t per = (int64_t)tiles_x * tiles_y * groups;
  while (zlen > 1 && per * ((cnnz + zlen - 1) / zlen) < 1024) zlen /= 2;
  const int zch = (cnnz + zlen - 1) / zlen;
  const int64_t nblk = per * zch;
  const T* dc = x0 ? (const T*)c.dinv : nullptr;
  if (w)
    hipLaunchKernelGGL((k_restrict222<T, NC, KC, true>), dim3((unsigned)nblk), dim3(256), 0, s,
                       (int)f.nx, (int)f.ny, (int)f.nz, tiles_x, tiles_y, zch, zlen, b, w, bc,
                       dc, cf, x0);
  else
    hipLaunchKernelGGL((k_restrict222<T, NC, KC, false>), dim3((unsigned)nblk), dim3(256), 0, s,
                       (int)f.nx, (int)f.ny, (int)f.nz, tiles_x, tiles_y, zch, zlen, b, w, bc,
                       dc, cf, x0);
  return true;
}

template <typename T, int NC, int KC>
static void launch_prolong(const MGLevel& f, const MGLevel& c, int groups, const T* xc,
                           const uint8_t* fixed, T* xf, hipStream_t s) {
  const RowGrid gp = row_grid(f.nx + 1, (f.ny + 1) * (f.nz + 1) * groups);
  if (fixed)
    hipLaunchKernelGGL((k_prolong_add<T, NC, KC, true>), gp.grid, gp.block, 0, s, (int)f.nx,
                       (int)f.ny, (int)f.nz, c.f[0], c.f[1], c.f[2], groups, xc, fixed, xf);
  else
    hipLaunchKernelGGL((k_prolong_add<T, NC, KC, false>), gp.grid, gp.block, 0, s, (int)f.nx,
                       (int)f.ny, (int)f.nz, c.f[0], c.f[1], c.f[2], groups, xc, fixed, xf);
}

template <typename T>
static const T* modal_coef(const skt_operator* op);
template <>
const double* modal_coef<double>(const skt_operator* op) { return op->modal_d; }
template <>
const float* modal_coef<float>(const skt_operator* op) { return op->modal_f; }

// V-cycle.  On the fine level the Chebyshev steps and the residual are fused
// into the matrix-free apply (hex_cheb_modal / hex_resid_modal); the
// iterate ping-pongs between z and lv[0].x, and the Chebyshev direction is
// the difference of the two buffers, so no direction vector is kept.  The
// number of fused steps, 2 * degree - 1, is odd: starting in lv[0].x the
// cycle ends in z.
template <typename T, int NC, int KC>
static int vcycle_t(MGData* mg, const skt_operator* op, const void* E, const void* dinv0, int k,
                    const T* r, T* z, T* w0, hipStream_t s, double* dot, int max_slots,
                    int* slots, const double* stop) {
  StOut<T> ost;
  ost.stop = stop;
  const int L = (int)mg->lv.size() - 1;
  const int groups = k / KC;
  const int deg = mg->degree;
  const uint8_t* fixed0 = op->n_fixed > 0 ? op->fixed : nullptr;
  const T* coef0 = modal_coef<T>(op);
  const T* E0 = (const T*)E;
  int stt;
  if (L >= 32) {
    set_error("internal: too many multigrid levels");
    return SKT_EINVAL;
  }
  // levels whose first smoothing iterate the restriction into them formed
  bool x0_done[32] = {};
  T* cur = (T*)mg->lv[0].x;  // fine iterate (ping-pong with z)
  T* oth = z;
  // Operators without the fused modal kernels (tet4 on a grid): the fine
  // level's Chebyshev steps as operator applies plus vector updates, with
  // the direction kept in z; the iterate lives in lv[0].x and is copied to z
  // at the end.
  const bool generic = !(op->mesh == SKT_MESH_HEX8 && op->modal_ok);
  if (generic) {
    MGLevel& v = mg->lv[0];
    const int64_t total = v.ndof * k;
    const unsigned ge = grid_for(total);
    const double* cf = mg->coef;
    T* x = (T*)v.x;
    T* d = z;
    for (int st = 0; st < deg; ++st) {
      T* w = st == 0 ? nullptr : w0;
      if (st > 0 && (stt = op_apply(const_cast<skt_operator*>(op), mg->dtype, k, KC, E, x, w0,
                                    s)))
        return stt;
      hipLaunchKernelGGL(k_cheb<T>, dim3(ge), dim3(kBlock), 0, s, total, v.ndof, KC,
                         cf + 2 * st, r, (const T*)w, (const T*)dinv0, d, x, st == 0 ? 1 : 0);
    }
    if ((stt = op_apply(const_cast<skt_operator*>(op), mg->dtype, k, KC, E, x, w0, s)))
      return stt;
    // zero at Dirichlet DOFs, as the fused restriction needs
    hipLaunchKernelGGL(k_rsub<T>, dim3(ge), dim3(kBlock), 0, s, total, v.ndof, KC, fixed0, r, w0);
    T* x01 = L > 1 ? (T*)mg->lv[1].x : nullptr;
    if (launch_restrict_fused<T, NC, KC>(v, mg->lv[1], groups, (const T*)w0, (const T*)nullptr,
                                         (T*)mg->lv[1].b, mg->coef + (size_t)2 * deg, x01, s))
      x0_done[1] = x01 != nullptr;
    else
      launch_restrict<T, NC, KC>(v, mg->lv[1], groups, (const T*)w0, (const T*)nullptr, fixed0,
                                 (T*)mg->lv[1].b, s);
    SKT_LAUNCH_CHECK();
    cur = x;
  } else {
    MGLevel& v = mg->lv[0];
    const int64_t total = v.ndof * k;
    const double* cf = mg->coef;
    if (deg == 2 && cheb0_enabled()) {
      // steps 0 and 1 in one pass (step 0's iterate is formed on the fly
      // and never stored, so this form serves degree 2 only: a third step
      // would need it as its previous iterate)
      const int ti = fine_timer_begin(mg, MG_T_CHEB0, s);
      if ((stt = hex_cheb0_modal<T>(op, k, KC, coef0, E0, r, oth, (const T*)dinv0, cf, cf + 2,
                                    s, stop)))
        return stt;
      fine_timer_end(mg, ti, s);
      std::swap(cur, oth);
    } else {
      hipLaunchKernelGGL(k_cheb_first<T>, dim3(grid_for(total)), dim3(kBlock), 0, s, total,
                         v.ndof, KC, cf, r, (const T*)dinv0, cur);
      // step st reads its previous iterate from the buffer it overwrites:
      // zero for step 1, the iterate of step st - 2 after that
      for (int st = 1; st < deg; ++st) {
        const int ti = fine_timer_begin(mg, MG_T_CHEB, s);
        if ((stt = hex_cheb_modal<T>(op, k, KC, coef0, E0, cur, oth, r, (const T*)dinv0,
                                     cf + 2 * st, st == 1 ? 1 : 0, s, nullptr, stop)))
          return stt;
        fine_timer_end(mg, ti, s);
        std::swap(cur, oth);
      }
    }
    const int tr = fine_timer_begin(mg, MG_T_RESID, s);
    if ((stt = hex_resid_modal<T>(op, k, KC, coef0, E0, cur, r, w0, s, stop))) return stt;
    fine_timer_end(mg, tr, s);
    // the residual is zero at Dirichlet DOFs, as the fused restriction needs;
    // it also forms level 1's first smoothing iterate unless level 1 is the
    // coarsest
    T* x01 = L > 1 ? (T*)mg->lv[1].x : nullptr;
    if (launch_restrict_fused<T, NC, KC>(v, mg->lv[1], groups, (const T*)w0, (const T*)nullptr,
                                         (T*)mg->lv[1].b, mg->coef + (size_t)2 * deg, x01, s))
      x0_done[1] = x01 != nullptr;
    else
      launch_restrict<T, NC, KC>(v, mg->lv[1], groups, (const T*)w0, (const T*)nullptr, fixed0,
                                 (T*)mg->lv[1].b, s);
    SKT_LAUNCH_CHECK();
  }
  // coarse levels: each level's iterate ping-pongs between its x and d
  // buffers through fused Chebyshev steps; it[l] holds the current one
  T* it[32];
  for (int l = 1; l < L; ++l) {
    MGLevel& v = mg->lv[l];
    const T* b = (const T*)v.b;
    T* A = (T*)v.x;
    T* B = (T*)v.d;
    T* w = (T*)v.w;
    const int64_t total = v.ndof * k;
    const double* cf = mg->coef + (size_t)l * 2 * deg;
    // step 0 from a zero iterate (formed by the fused restriction into this
    // level, or here), then fused steps (the iterate before step 0 is zero)
    if (!x0_done[l])
      hipLaunchKernelGGL(k_cheb_first<T>, dim3(grid_for(total)), dim3(kBlock), 0, s, total,
                         v.ndof, KC, cf, b, (const T*)v.dinv, A);
    for (int st = 1; st < deg; ++st) {
      if ((stt = level_cheb<T>(mg, op, l, k, KC, A, B, cf + 2 * st, st == 1 ? 1 : 0, s, stop)))
        return stt;
      std::swap(A, B);
    }
    it[l] = A;
    if ((stt = level_apply<T>(mg, op, E, l, k, KC, A, w, s, ost))) return stt;
    T* x0n = l + 1 < L ? (T*)mg->lv[l + 1].x : nullptr;
    if (launch_restrict_fused<T, NC, KC>(v, mg->lv[l + 1], groups, b, w, (T*)mg->lv[l + 1].b,
                                         mg->coef + (size_t)(l + 1) * 2 * deg, x0n, s))
      x0_done[l + 1] = x0n != nullptr;
    else
      launch_restrict<T, NC, KC>(v, mg->lv[l + 1], groups, b, w, nullptr, (T*)mg->lv[l + 1].b,
                                 s);
    SKT_LAUNCH_CHECK();
  }
  // coarsest
  {
    const MGLevel& c = mg->lv[L];
    launch_dense_apply<T, KC>(c.ndof, groups, (const T*)mg->Ainv, (const T*)c.b, (T*)c.x, s);
    SKT_LAUNCH_CHECK();
    it[L] = (T*)c.x;
  }
  // coarse levels, upward
  for (int l = L - 1; l >= 1; --l) {
    MGLevel& v = mg->lv[l];
    T* A = it[l];
    T* B = (A == (T*)v.x) ? (T*)v.d : (T*)v.x;
    const double* cf = mg->coef + (size_t)l * 2 * deg;
    launch_prolong<T, NC, KC>(v, mg->lv[l + 1], groups, (const T*)it[l + 1], nullptr, A, s);
    for (int st = 0; st < deg; ++st) {
      // step 0 has c1 = 0 and needs no previous iterate
      if ((stt = level_cheb<T>(mg, op, l, k, KC, A, B, cf + 2 * st, st == 0 ? 1 : 0, s, stop)))
        return stt;
      std::swap(A, B);
    }
    it[l] = A;
    SKT_LAUNCH_CHECK();
  }
  if (generic) {
    MGLevel& v = mg->lv[0];
    const int64_t total = v.ndof * k;
    const unsigned ge = grid_for(total);
    const double* cf = mg->coef;
    T* x = (T*)v.x;
    launch_prolong<T, NC, KC>(v, mg->lv[1], groups, (const T*)it[1], fixed0, x, s);
    // post-smoothing from the corrected iterate (step 0 has c1 = 0, so the
    // stale direction in z is not used)
    for (int st = 0; st < deg; ++st) {
      if ((stt = op_apply(const_cast<skt_operator*>(op), mg->dtype, k, KC, E, x, w0, s)))
        return stt;
      hipLaunchKernelGGL(k_cheb<T>, dim3(ge), dim3(kBlock), 0, s, total, v.ndof, KC,
                         cf + 2 * st, r, (const T*)w0, (const T*)dinv0, z, x, 0);
    }
    hipLaunchKernelGGL(k_copy_vec<T>, dim3(ge), dim3(kBlock), 0, s, total, (const T*)x, z);
    if (slots) *slots = 0;
    SKT_LAUNCH_CHECK();
    return SKT_OK;
  }
  // fine level, correction and post-smoothing
  {
    MGLevel& v = mg->lv[0];
    const double* cf = mg->coef;
    launch_prolong<T, NC, KC>(v, mg->lv[1], groups, (const T*)it[1], fixed0, cur, s);
    // the last step can leave the partials of r . z
    const int ns = hex_modal_dot_slots(op, k, KC);
    double* last_dot = (dot && slots && ns <= max_slots) ? dot : nullptr;
    for (int st = 0; st < deg; ++st) {
      // step 0 has c1 = 0 and needs no previous iterate
      double* sd = st == deg - 1 ? last_dot : nullptr;
      const int ti = fine_timer_begin(mg, sd ? MG_T_CHEB_DOT : MG_T_CHEB, s);
      if ((stt = hex_cheb_modal<T>(op, k, KC, coef0, E0, cur, oth, r, (const T*)dinv0,
                                   cf + 2 * st, 0, s, sd, stop)))
        return stt;
      fine_timer_end(mg, ti, s);
      std::swap(cur, oth);
    }
    if (slots) *slots = (last_dot && deg >= 1) ? ns : 0;
    SKT_LAUNCH_CHECK();
  }
  if (cur != z) {
    set_error("internal: multigrid iterate did not end in z");
    return SKT_EINVAL;
  }
  return SKT_OK;
}

int mg_vcycle(MGData* mg, const skt_operator* op, const void* E, const void* dinv0, int ncols,
              int kc, const void* r, void* z, void* w0, hipStream_t s, double* dot,
              int max_slots, int* slots, const double* stop) {
  SKT_CHECK(mg->dtype >= 0, "multigrid not set up");
  SKT_CHECK(ncols <= mg->max_cols && ncols % kc == 0, "bad column count for multigrid");
#define SKT_VC(T_, NC_, KC_)                                                               \
  return vcycle_t<T_, NC_, KC_>(mg, op, E, dinv0, ncols, (const T_*)r, (T_*)z, (T_*)w0, s, dot, \
                                max_slots, slots, stop)
  if (mg->dtype == SKT_F64) {
    if (op->ncomp == 3) {
      if (kc == 1) SKT_VC(double, 3, 1);
      if (kc == 2) SKT_VC(double, 3, 2);
    } else {
      if (kc == 1) SKT_VC(double, 1, 1);
      if (kc == 2) SKT_VC(double, 1, 2);
      if (kc == 4) SKT_VC(double, 1, 4);
    }
  } else {
    if (op->ncomp == 3) {
      if (kc == 1) SKT_VC(float, 3, 1);
      if (kc == 2) SKT_VC(float, 3, 2);
    } else {
      if (kc == 1) SKT_VC(float, 1, 1);
      if (kc == 2) SKT_VC(float, 1, 2);
      if (kc == 4) SKT_VC(float, 1, 4);
    }
  }
#undef SKT_VC
  set_error("unsupported column block for multigrid");
  return SKT_EINVAL;
}

}  // namespace skt
