// Element <-> node transfers of the nodal Helmholtz filter on structured hex8
// grids and tet4 meshes (filters/helmholtz_filter_nodal.py:27-59).
//
// Element -> node: volume-weighted average over the (up to 8) elements of a
// node, non-design elements counting with a fixed value.  The reference
// accumulates element by element in ascending element index; the gather
// below visits a node's elements in the same ascending order, so the sums
// (and the quotient) are the same floating-point operations.
//
// Node -> element: mean of the 8 corner values, summed in the mesh's local
// node order and divided by 8, as numpy.mean over the node axis does.
// Both transfers work on ncols columns at once ((n, ncols) row-major); the
// node -> element one can also fold the columns into their mean, summed in
// column order from 0.0 like the reference's per-load-case accumulation.
#include "operator.hpp"

namespace skt {

// skfem local node order of a hex (matches the operators' corner order)
__constant__ int kNx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ int kNy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__constant__ int kNz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

__global__ void k_elem_to_node(int64_t nx, int64_t ny, int64_t nz, int ncols,
                               const double* __restrict__ vol, const uint8_t* __restrict__ design,
                               double fixed_value, const double* __restrict__ re,
                               double* __restrict__ rn) {
  const int64_t nnx = nx + 1, nny = ny + 1;
  const int64_t total = nnx * nny * (nz + 1) * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = t / ncols;
    const int col = (int)(t - node * ncols);
    const int64_t i = node % nnx, j = (node / nnx) % nny, k = node / (nnx * nny);
    double acc = 0.0, ws = 0.0;
    // slot s = sx + 2 sy + 4 sz in ascending order visits ascending elements
    for (int s = 0; s < 8; ++s) {
      const int64_t ex = i - 1 + (s & 1), ey = j - 1 + ((s >> 1) & 1), ez = k - 1 + (s >> 2);
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
      const int64_t e = ex + nx * (ey + ny * ez);
      const double w = vol[e];
      const double v = (!design || design[e]) ? re[e * ncols + col] : fixed_value;
      acc += w * v;
      ws += w;
    }
    rn[t] = acc / (ws == 0.0 ? 1.0 : ws);
  }
}

__global__ void k_node_to_elem(int64_t nx, int64_t ny, int64_t nz, int ncols,
                               const double* __restrict__ rn, int clamp_nonpositive,
                               int mean_cols, double* __restrict__ re) {
  const int64_t nnx = nx + 1, nny = ny + 1;
  const int64_t ne = nx * ny * nz;
  const int64_t total = mean_cols ? ne : ne * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = mean_cols ? t : t / ncols;
    const int c0 = mean_cols ? 0 : (int)(t - e * ncols);
    const int c1 = mean_cols ? ncols : c0 + 1;
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    double sum = 0.0, v = 0.0;
    for (int c = c0; c < c1; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 8; ++a)
        acc += rn[((ex + kNx[a]) + nnx * ((ey + kNy[a]) + nny * (ez + kNz[a]))) * ncols + c];
      v = acc / 8.0;
      if (clamp_nonpositive && v > 0.0) v = 0.0;
      sum += v;   // the accumulation starts from 0.0 as in the reference
    }
    re[t] = mean_cols ? sum / (double)ncols : v;
  }
}

// Tet4 element -> node: the node's elements come from the operator's
// node -> element lists, which are sorted by element index, so the
// accumulation visits them in the reference's (ascending) order.
__global__ void k_tet_elem_to_node(int64_t n_nodes, int ncols, const int64_t* __restrict__ nptr,
                                   const int32_t* __restrict__ nel,
                                   const double* __restrict__ vol,
                                   const uint8_t* __restrict__ design, double fixed_value,
                                   const double* __restrict__ re, double* __restrict__ rn) {
  const int64_t total = n_nodes * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = t / ncols;
    const int col = (int)(t - node * ncols);
    double acc = 0.0, ws = 0.0;
    for (int64_t q = nptr[node]; q < nptr[node + 1]; ++q) {
      const int64_t e = nel[q];
      const double w = vol[e];
      const double v = (!design || design[e]) ? re[e * ncols + col] : fixed_value;
      acc += w * v;
      ws += w;
    }
    rn[t] = acc / (ws == 0.0 ? 1.0 : ws);
  }
}

// Tet4 node -> element: mean of the 4 vertex values in the mesh's local
// order, divided by 4 (numpy.mean over the node axis).
__global__ void k_tet_node_to_elem(int64_t n_el, int ncols, const int32_t* __restrict__ conn,
                                   const double* __restrict__ rn, int clamp_nonpositive,
                                   int mean_cols, double* __restrict__ re) {
  const int64_t total = mean_cols ? n_el : n_el * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = mean_cols ? t : t / ncols;
    const int c0 = mean_cols ? 0 : (int)(t - e * ncols);
    const int c1 = mean_cols ? ncols : c0 + 1;
    double sum = 0.0, v = 0.0;
    for (int c = c0; c < c1; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) acc += rn[(int64_t)conn[a * n_el + e] * ncols + c];
      v = acc / 4.0;
      if (clamp_nonpositive && v > 0.0) v = 0.0;
      sum += v;
    }
    re[t] = mean_cols ? sum / (double)ncols : v;
  }
}

}  // namespace skt

using namespace skt;

extern "C" {

int skt_hex8_elem_to_node(int64_t nx, int64_t ny, int64_t nz, int ncols, const double* vol,
                          const uint8_t* design, double fixed_value, const double* rho_elem,
                          double* rho_node, hipStream_t stream) {
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(ncols >= 1, "ncols must be >= 1");
  SKT_CHECK(vol && rho_elem && rho_node, "null pointer");
  const int64_t total = (nx + 1) * (ny + 1) * (nz + 1) * ncols;
  hipLaunchKernelGGL(k_elem_to_node, dim3(grid_for(total)), dim3(kBlock), 0, stream, nx, ny, nz,
                     ncols, vol, design, fixed_value, rho_elem, rho_node);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_hex8_node_to_elem(int64_t nx, int64_t ny, int64_t nz, int ncols, const double* rho_node,
                          int clamp_nonpositive, int mean_cols, double* rho_elem,
                          hipStream_t stream) {
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be >= 1");
  SKT_CHECK(ncols >= 1, "ncols must be >= 1");
  SKT_CHECK(rho_node && rho_elem, "null pointer");
  const int64_t ne = nx * ny * nz;
  const int64_t total = mean_cols ? ne : ne * ncols;
  hipLaunchKernelGGL(k_node_to_elem, dim3(grid_for(total)), dim3(kBlock), 0, stream, nx, ny, nz,
                     ncols, rho_node, clamp_nonpositive, mean_cols, rho_elem);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_tet4_elem_to_node(const skt_operator* op, int ncols, const double* vol,
                          const uint8_t* design, double fixed_value, const double* rho_elem,
                          double* rho_node, hipStream_t stream) {
  SKT_CHECK(op && op->mesh == SKT_MESH_TET4, "a tet4 operator is required");
  SKT_CHECK(ncols >= 1, "ncols must be >= 1");
  SKT_CHECK(vol && rho_elem && rho_node, "null pointer");
  const int64_t total = op->n_nodes * ncols;
  hipLaunchKernelGGL(k_tet_elem_to_node, dim3(grid_for(total)), dim3(kBlock), 0, stream,
                     op->n_nodes, ncols, op->nptr, op->nel, vol, design, fixed_value, rho_elem,
                     rho_node);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_tet4_node_to_elem(const skt_operator* op, int ncols, const double* rho_node,
                          int clamp_nonpositive, int mean_cols, double* rho_elem,
                          hipStream_t stream) {
  SKT_CHECK(op && op->mesh == SKT_MESH_TET4, "a tet4 operator is required");
  SKT_CHECK(ncols >= 1, "ncols must be >= 1");
  SKT_CHECK(rho_node && rho_elem, "null pointer");
  const int64_t total = mean_cols ? op->n_el : op->n_el * ncols;
  hipLaunchKernelGGL(k_tet_node_to_elem, dim3(grid_for(total)), dim3(kBlock), 0, stream,
                     op->n_el, ncols, op->conn, rho_node, clamp_nonpositive, mean_cols,
                     rho_elem);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

}  // extern "C"
