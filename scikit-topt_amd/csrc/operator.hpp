// Internal view of the matrix-free operators shared by operators.hip and
// solver.hip.
#pragma once

#include "common.hpp"

enum { SKT_MESH_HEX8 = 0, SKT_MESH_TET4 = 1 };

struct skt_operator {
  int mesh = SKT_MESH_HEX8;
  int physics = SKT_ELASTIC;
  int ncomp = 3;
  int64_t n_nodes = 0, n_el = 0, ndof = 0;
  // hex8 structured grid
  int64_t nx = 0, ny = 0, nz = 0;
  double h[3] = {1, 1, 1};
  double* K0d = nullptr;   // (8 ncomp)^2 float64, unit modulus
  float* K0f = nullptr;    // float32 copy
  double K0h[24 * 24];     // host copy
  // tet4
  int32_t* conn = nullptr;     // (4, n_el)
  double* g64 = nullptr;       // (12, n_el) barycentric gradients * sqrt-free
  float* g32 = nullptr;
  double* vol64 = nullptr;     // (n_el)
  float* vol32 = nullptr;
  int64_t* nptr = nullptr;     // (n_nodes + 1) node -> element CSR
  int32_t* nel = nullptr;
  uint8_t* nloc = nullptr;
  double lam0 = 0.0, mu0 = 0.0;  // unit-modulus Lame parameters
  // Dirichlet DOFs
  uint8_t* fixed = nullptr;    // (ndof)
};

namespace skt {
int op_apply(skt_operator* op, int dtype, int ncols, const void* E,
             const void* X, void* Y, hipStream_t stream);
int op_diag(skt_operator* op, int dtype, const void* E, void* diag,
            hipStream_t stream);
}  // namespace skt
