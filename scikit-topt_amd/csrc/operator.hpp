// Internal view of the matrix-free operators shared by operators.hip and
// solver.hip.
#pragma once

#include "common.hpp"

enum { SKT_MESH_HEX8 = 0, SKT_MESH_TET4 = 1 };

// Column stop flags: column c is stopped where stop[c * kStopStride] != 0
// (the solver's per-column scalar block).  Launches given such flags may
// skip the work of workgroups whose columns have all stopped; their
// outputs (and dot partials) for those columns are then left stale.
constexpr int kStopStride = 16;

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
  // modal (Walsh-Hadamard) coefficients of K_e^0 when it has the box pattern
  int modal_ok = 0;
  double modal_d[40];
  float modal_f[40];
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
  // scalar tet4 operators: a M + b K (mass and conduction), a = 0 for plain
  // conduction (skt_tet4_create_scalar sets them)
  double mass_a = 0.0, cond_b = 1.0;
  // 1: the tet4 mesh is a grid of nx*ny*nz cells, six tets per cell
  // (skt_tet4_set_grid), which lets the multigrid coarsen it
  int tet_grid = 0;
  // with tet_grid: per tet, the cell corner (bits x, y, z) of each local node
  // q in bits 3q..3q+2
  int32_t* tet_code = nullptr;
  // tet4 apply staging: ncomp nodal forces per (element, corner) pair and
  // column, stored in the node -> element lists' order; tet_slot[a n_el + e]
  // is the position of (e, a) in those lists
  void* tet_f = nullptr;
  size_t tet_f_bytes = 0;
  int32_t* tet_slot = nullptr;
  // scalar tet4 operators: extra symmetric 4 x 4 element matrices (Robin
  // terms) added on free rows, packed (10, n_el) in the order 00 11 22 33 01
  // 02 03 12 13 23 (skt_tet4_set_extra), or null
  double* tet_x64 = nullptr;
  float* tet_x32 = nullptr;
  // extra scalar 27-point stencil added on free rows (Robin terms), or null
  double* extra64 = nullptr;   // (27, n_nodes)
  float* extra32 = nullptr;
  // Dirichlet DOFs
  uint8_t* fixed = nullptr;    // (ndof)
  int32_t* fixed_list = nullptr;  // indices of the fixed DOFs
  int64_t n_fixed = 0;
};

namespace skt {
// Y += S X on free rows for the operator's extra stencil (robin.hip); no-op
// without one.  diag += the stencil's centre entries.
int extra_apply_add(const skt_operator* op, int dtype, int ncols, int kc, const void* X, void* Y,
                    hipStream_t s);
int extra_diag_add(const skt_operator* op, int dtype, void* diag, hipStream_t s);
// Shared tail of the hex8 constructors (operators.hip).
int hex8_finish(skt_operator* op, const uint8_t* dof_fixed, skt_operator** out);
// Element energies 0.5 E_e u_e^T K0 u_e ((n_el, ncols) row-major U and out)
// through the modal form of K0.
template <typename T>
int hex_energy_modal(const skt_operator* op, int ncols, const T* coef, const T* E, const T* U,
                     double* out, hipStream_t s);
// Output modes of the structured hex8 modal kernel.
// OUT_CHEB0: the first two Chebyshev steps from a zero iterate in one pass:
// X is the right-hand side b, the first iterate x1 = s0 dinv b is formed on
// the fly (s0 = cheb0[1]) and x_out = x1 + c1 x1 + c2 dinv (b - K x1).
enum { OUT_APPLY = 0, OUT_RESID = 1, OUT_CHEB = 2, OUT_CHEB0 = 3 };
// Partial sums per column left by a fused dot product (hex_apply_modal and
// hex_cheb_modal with `dot`): dot[c * slots + i], i < slots, for column c.
int hex_modal_dot_slots(const skt_operator* op, int k, int kc);
// Y = K X; with `dot`, also the partials of X . Y per column
template <typename T>
int hex_apply_modal(const skt_operator* op, int k, int kc, const T* coef, const T* E, const T* X,
                    T* Y, hipStream_t s, double* dot = nullptr, const double* stop = nullptr);
// R = B - K X, zero at Dirichlet DOFs
template <typename T>
int hex_resid_modal(const skt_operator* op, int k, int kc, const T* coef, const T* E,
                    const T* X, const T* B, T* R, hipStream_t s, const double* stop = nullptr);
// x_out = x1 + c1 x1 + c2 dinv (b - K x1) with x1 = cheb0[1] dinv b (OUT_CHEB0)
template <typename T>
int hex_cheb0_modal(const skt_operator* op, int k, int kc, const T* coef, const T* E, const T* B,
                    T* Xout, const T* dinv, const double* cheb0, const double* cheb,
                    hipStream_t s, const double* stop = nullptr);
// Chebyshev-Jacobi step with the direction taken from two iterates:
// Xout = X + c1 (X - Xprev) + c2 dinv (B - K X), (c1, c2) = cheb[0..1] on the
// device.  Xprev is read from Xout before it is overwritten (zero when
// prev_zero); Xout must not alias X.  With `dot`, also the partials of
// B . Xout per column.
template <typename T>
int hex_cheb_modal(const skt_operator* op, int k, int kc, const T* coef, const T* E, const T* X,
                   T* Xout, const T* B, const T* dinv, const double* cheb, int prev_zero,
                   hipStream_t s, double* dot = nullptr, const double* stop = nullptr);
// Y = K(E) X for ncols columns.  kc == ncols: (ndof, ncols) row-major;
// otherwise the column-blocked layout [ncols/kc][ndof][kc] (structured hex8
// only).
// Timing of the next modal-kernel launch on this host thread: the launch
// records `start` and `stop` itself (hipExtLaunchKernelGGL).  modal_time_clear
// drops a request no launch took and returns whether one was pending.
void modal_time_next(hipEvent_t start, hipEvent_t stop);
bool modal_time_clear();

int op_apply(skt_operator* op, int dtype, int ncols, int kc, const void* E,
             const void* X, void* Y, hipStream_t stream);
// op_apply that also leaves the per-column partials of X . Y in dot (see
// hex_modal_dot_slots) where the operator can fuse them: *slots is then
// their count per column, else 0 and nothing is written to dot.
int op_apply_dot(skt_operator* op, int dtype, int ncols, int kc, const void* E, const void* X,
                 void* Y, double* dot, int max_slots, int* slots, hipStream_t stream,
                 const double* stop = nullptr);
int op_diag(skt_operator* op, int dtype, const void* E, void* diag,
            hipStream_t stream);
}  // namespace skt
