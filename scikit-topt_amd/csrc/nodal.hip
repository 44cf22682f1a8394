// Element <-> node transfers of the nodal Helmholtz filter on structured hex8
// grids (filters/helmholtz_filter_nodal.py:27-59).
//
// Element -> node: volume-weighted average over the (up to 8) elements of a
// node, non-design elements counting with a fixed value.  The reference
// accumulates element by element in ascending element index; the gather
// below visits a node's elements in the same ascending order, so the sums
// (and the quotient) are the same floating-point operations.
//
// Node -> element: mean of the 8 corner values, summed in the mesh's local
// node order and divided by 8, as numpy.mean over the node axis does.
#include "common.hpp"

namespace skt {

// skfem local node order of a hex (matches the operators' corner order)
__constant__ int kNx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ int kNy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__constant__ int kNz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

__global__ void k_elem_to_node(int64_t nx, int64_t ny, int64_t nz, const double* __restrict__ vol,
                               const uint8_t* __restrict__ design, double fixed_value,
                               const double* __restrict__ re, double* __restrict__ rn) {
  const int64_t nnx = nx + 1, nny = ny + 1;
  const int64_t nn = nnx * nny * (nz + 1);
  for (int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; node < nn;
       node += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = node % nnx, j = (node / nnx) % nny, k = node / (nnx * nny);
    double acc = 0.0, ws = 0.0;
    // slot s = sx + 2 sy + 4 sz in ascending order visits ascending elements
    for (int s = 0; s < 8; ++s) {
      const int64_t ex = i - 1 + (s & 1), ey = j - 1 + ((s >> 1) & 1), ez = k - 1 + (s >> 2);
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
      const int64_t e = ex + nx * (ey + ny * ez);
      const double w = vol[e];
      const double v = (!design || design[e]) ? re[e] : fixed_value;
      acc += w * v;
      ws += w;
    }
    rn[node] = acc / (ws == 0.0 ? 1.0 : ws);
  }
}

__global__ void k_node_to_elem(int64_t nx, int64_t ny, int64_t nz, const double* __restrict__ rn,
                               int clamp_nonpositive, double* __restrict__ re) {
  const int64_t nnx = nx + 1, nny = ny + 1;
  const int64_t ne = nx * ny * nz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a)
      acc += rn[(ex + kNx[a]) + nnx * ((ey + kNy[a]) + nny * (ez + kNz[a]))];
    double v = acc / 8.0;
    if (clamp_nonpositive && v > 0.0) v = 0.0;
    re[e] = v;
  }
}

}  // namespace skt

using namespace skt;

extern "C" {

int skt_hex8_elem_to_node(int64_t nx, int64_t ny, int64_t nz, const double* vol,
                          const uint8_t* design, double fixed_value, const double* rho_elem,
                          double* rho_node, hipStream_t stream) {
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(vol && rho_elem && rho_node, "null pointer");
  const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL(k_elem_to_node, dim3(grid_for(nn)), dim3(kBlock), 0, stream, nx, ny, nz, vol,
                     design, fixed_value, rho_elem, rho_node);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_hex8_node_to_elem(int64_t nx, int64_t ny, int64_t nz, const double* rho_node,
                          int clamp_nonpositive, double* rho_elem, hipStream_t stream) {
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(rho_node && rho_elem, "null pointer");
  const int64_t ne = nx * ny * nz;
  hipLaunchKernelGGL(k_node_to_elem, dim3(grid_for(ne)), dim3(kBlock), 0, stream, nx, ny, nz,
                     rho_node, clamp_nonpositive, rho_elem);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

}  // extern "C"
