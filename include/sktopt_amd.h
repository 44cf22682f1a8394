/*
 * sktopt_amd -- C ABI of the MI355X (gfx950) density-method hot path.
 *
 * Every entry point takes device pointers (owned by the caller, never freed
 * by the library), plain sizes and a hipStream_t, enqueues its work on that
 * stream and returns an int status (SKT_OK on success).  skt_last_error()
 * returns a message for the last failing call on the calling host thread.
 * Handles (skt_operator, skt_filter, skt_solver) own their internal
 * workspaces and release them in their destroy call.
 *
 * Reference interfaces replaced (paths relative to scikit-topt/sktopt/):
 * see the comment above each declaration.  The reference is pure Python;
 * the binding a maintainer would add is a ctypes stub.
 *
 * Array conventions
 *   element arrays      (n_el,)                 float64 unless stated
 *   block vectors       (ndof, ncols) row-major, i.e. the ncols load cases of
 *                       one DOF are contiguous; ndof = ncomp * n_nodes and
 *                       dof = ncomp * node + component.
 *   dtype codes         SKT_F64 = 0, SKT_F32 = 1
 */
#ifndef SKTOPT_AMD_H
#define SKTOPT_AMD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SKT_OK 0
#define SKT_EINVAL 1      /* bad argument (maps to ValueError)                 */
#define SKT_EHIP 2        /* HIP runtime error (maps to RuntimeError)          */
#define SKT_ENOCONV 3     /* iterative solve did not reach its tolerance        */
#define SKT_ENOTIMPL 4    /* unsupported combination (NotImplementedError)      */

#define SKT_F64 0
#define SKT_F32 1

#define SKT_SIMP 0
#define SKT_RAMP 1
/* RAMP moduli with the exact derivative (1 + p) / (1 + p (1 - rho))^2 for
 * the dC kernels; core/derivatives.py:56-61 (SKT_RAMP) uses
 * (den - p rho) / den^2 instead.  No reference counterpart: it serves the
 * differentiable compliance, whose gradient must match finite differences. */
#define SKT_RAMP_EXACT 2

#define SKT_ELASTIC 0
#define SKT_HEAT 1

typedef struct skt_operator skt_operator;
typedef struct skt_filter skt_filter;
typedef struct skt_solver skt_solver;

const char* skt_last_error(void);
int skt_abi_version(void);

/* ------------------------------------------------------------------------
 * Design-variable transforms (float64, elementwise)
 * ---------------------------------------------------------------------- */

/* core/projection.py:27-77  heaviside_projection_inplace(x, beta, eta, out) */
int skt_heaviside_project(int64_t n, const double* x, double beta, double eta,
                          double* out, hipStream_t stream);

/* core/projection.py:80-118  heaviside_projection_derivative_inplace */
int skt_heaviside_derivative(int64_t n, const double* x, double beta,
                             double eta, double* out, hipStream_t stream);

/* fea/composer.py:19-39  simp_interpolation / ramp_interpolation.
 * out has dtype out_dtype (SKT_F64 or SKT_F32). */
int skt_interpolate(int kind, int64_t n, const double* rho, double E0,
                    double Emin, double p, void* out, int out_dtype,
                    hipStream_t stream);

/* core/derivatives.py:42-52,65-68  dC_drho_simp / dC_drho_ramp */
int skt_dC_drho(int kind, int64_t n, const double* rho, const double* energy,
                int64_t energy_stride, double E0, double Emin, double p,
                double* out, hipStream_t stream);

/* core/optimizers/common_density.py:1073-1101 fused over load cases:
 * g[e] = scale * sum_k dC_kind(rho_proj[e], energy[e,k]) * H'(rho_filt[e]).
 * energy is (n_el, ncols) row-major; scale = 1/n_tasks folds the mean
 * (the filter adjoint that follows is linear). */
int skt_sensitivity_chain(int kind, int64_t n, int ncols,
                          const double* rho_proj, const double* energy,
                          const double* rho_filt, double beta, double eta,
                          double E0, double Emin, double p, double scale,
                          double* g, hipStream_t stream);

/* The same chain per load case, g (n, ncols) row-major, without the mean:
 * g[:, k] = dC(rho_proj, energy[:, k]) * H'(rho_filt)
 * (core/optimizers/common_density.py:1075-1097), for filters whose adjoint
 * is not linear (the nodal Helmholtz gradient clamps each call). */
int skt_sensitivity_chain_cols(int kind, int64_t n, int ncols,
                               const double* rho_proj, const double* energy,
                               const double* rho_filt, double beta, double eta,
                               double E0, double Emin, double p, double* g,
                               hipStream_t stream);

/* Both chains with a per-load-case additive term (nullable, (n, ncols)):
 * dC_k = dC(rho_proj, energy[:, k]) + extra[:, k] before the product with
 * H' -- the custom sensitivity of the Robin heat compliance
 * (fea/solver_heat.py:928-980, common_density.py:1067-1097).  per_column = 0:
 * g (n,) = scale * sum_k dC_k H'; per_column = 1: g (n, ncols) = dC_k H'. */
int skt_sensitivity_chain_extra(int kind, int64_t n, int ncols,
                                const double* rho_proj, const double* energy,
                                const double* extra, const double* rho_filt,
                                double beta, double eta, double E0, double Emin,
                                double p, double scale, int per_column,
                                double* g, hipStream_t stream);

/* Elementwise helpers of the update rules (logmoc.py:186-225):
 * op 0: out = a*x + b*y (y may be NULL: out = a*x); op 1: out = a*x*y. */
int skt_ewise(int op, int64_t n, double a, const double* x, double b,
              const double* y, double* out, hipStream_t stream);

/* ------------------------------------------------------------------------
 * Reductions (float64; results are written to device memory).  The
 * reductions, statistics and percentile keep their partials in a buffer of
 * the calling (device, stream) pair, allocated on the stream's first use:
 * calls on one stream are ordered, and calls on different streams use
 * different buffers.
 * ---------------------------------------------------------------------- */
#define SKT_RED_SUM 0        /* sum x                                     */
#define SKT_RED_DOT 1        /* sum x*y                                   */
#define SKT_RED_MAXABS 2     /* max |x|                                   */
#define SKT_RED_MAXABSDIFF 3 /* max |x - y|                               */

int skt_reduce(int op, int64_t n, const double* x, const double* y,
               double* out, hipStream_t stream);

/* Column-wise reductions of a (n, ncols) row-major block:
 * op SKT_RED_SUM -> out[c] = sum_i x[i,c];
 * op SKT_RED_DOT -> out[c] = sum_i x[i,c]*y[i,c];
 * op SKT_RED_MAXABS -> out[c] = max_i |x[i,c]|. */
int skt_reduce_cols(int op, int64_t n, int ncols, const double* x,
                    const double* y, double* out, hipStream_t stream);

/* out[i] = mean of row i of a (n, k) row-major block, k <= 128, summed in
 * numpy's pairwise order for short rows: energy.mean(axis=1) of
 * core/optimizers/common_density.py:1065, recorded as "energy" (:1148). */
int skt_row_mean(int64_t n, int k, const double* x, double* out, hipStream_t stream);

/* out[0..3] = [min, mean, max, std] of x (population std, numpy's np.std),
 * the statistics a "min-max-mean-std" history keeps per iteration
 * (tools/history.py:145-151).  n >= 1. */
int skt_stats4(int64_t n, const double* x, double* out, hipStream_t stream);

/* np.percentile(np.abs(x), q) with numpy's default 'linear' method
 * (core/optimizers/oc.py:188-189, logmoc.py:167-174,220).  Exact radix
 * select on the float64 bit patterns. */
int skt_percentile_abs(int64_t n, const double* x, double q, double* out,
                       hipStream_t stream);

/* skt_percentile_abs without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * graph capture, tight serving loops).  Its scratch is a static 17.5 KB
 * device array per GPU, held until the process ends; n <= 2^53.  Every call
 * on one GPU shares that array, so calls that overlap on one GPU (two
 * streams or two host threads) race. */
int skt_percentile_abs_noalloc(int64_t n, const double* x, double q, double* out,
                               hipStream_t stream);

/* ------------------------------------------------------------------------
 * Matrix-free SIMP/RAMP-weighted operators
 *   K(E) = sum_e E_e K_e^0, Dirichlet rows replaced by identity
 *   (skfem.enforce semantics as used at fea/solver_elastic.py:205-211).
 * ---------------------------------------------------------------------- */

/* Structured nx*ny*nz box grid of hx*hy*hz hexahedra (hex8), nodes numbered
 * x-fastest.  physics SKT_ELASTIC (3 dof/node, Poisson ratio nu) or SKT_HEAT
 * (1 dof/node).  quad_order: Gauss points per direction for K_e^0 (>= 1).
 * dof_fixed: (ndof,) uint8 device array, 1 = Dirichlet DOF (copied).
 * Replaces composer.assemble_stiffness_matrix (fea/composer.py:53-101) and
 * assemble_conduction_matrix (fea/composer.py:104-145) on hex meshes. */
int skt_hex8_create(int physics, int64_t nx, int64_t ny, int64_t nz,
                    double hx, double hy, double hz, double nu,
                    int quad_order, const uint8_t* dof_fixed,
                    skt_operator** out);

/* Structured hex8 scalar operator a M + b K (M: unit-density trilinear mass,
 * K: unit-conductivity Laplacian), same grid conventions as skt_hex8_create.
 * With a = 1, b = r^2 it is the nodal Helmholtz filter's bilinear form
 * u v + r^2 grad u . grad v (filters/helmholtz_filter_nodal.py:140-142);
 * with b = 0 the mass matrix of its right-hand side (:144-146). */
int skt_hex8_create_scalar(int64_t nx, int64_t ny, int64_t nz, double hx,
                           double hy, double hz, double mass_coef,
                           double stiff_coef, const uint8_t* dof_fixed,
                           skt_operator** out);

/* filters/helmholtz_filter_nodal.py:31-59 element_to_node_density_averaging:
 * rho_node = sum_e w_e v_e / sum_e w_e over the elements of each node,
 * w = vol, v_e = rho_elem[e] on design elements (design[e] != 0, or all
 * when design is NULL) and fixed_value elsewhere; same summation order as
 * the reference's loop.  rho_elem (n_el, ncols) -> rho_node (n_nodes, ncols),
 * row-major, one averaging per column. */
int skt_hex8_elem_to_node(int64_t nx, int64_t ny, int64_t nz, int ncols,
                          const double* vol, const uint8_t* design,
                          double fixed_value, const double* rho_elem,
                          double* rho_node, hipStream_t stream);

/* helmholtz_filter_nodal.py:27-28 node_to_element_density (mean of the 8
 * nodes) per column of rho_node (n_nodes, ncols); clamp_nonpositive applies
 * the gradient's min(v, 0) (:232).  mean_cols = 0: rho_elem is
 * (n_el, ncols); mean_cols = 1: rho_elem is (n_el,) = the columns summed in
 * order and divided by ncols, as the per-load-case accumulation of the
 * filtered sensitivities (core/optimizers/common_density.py:1098, :1101). */
int skt_hex8_node_to_elem(int64_t nx, int64_t ny, int64_t nz, int ncols,
                          const double* rho_node, int clamp_nonpositive,
                          int mean_cols, double* rho_elem, hipStream_t stream);

/* Tet4 scalar operator a M + b K (M: P1 mass V/20 (1 + delta_ab), K: unit-
 * conductivity Laplacian), the nodal Helmholtz filter's bilinear form on
 * tetrahedral meshes (filters/helmholtz_filter_nodal.py:140-146 with
 * ElementTetP1, whose default quadrature integrates both exactly).
 * node_fixed: (n_nodes,) uint8, identity rows.  Energies of this operator
 * are those of its conduction part. */
int skt_tet4_create_scalar(int64_t n_nodes, int64_t n_el, const double* coords,
                           const int32_t* conn, double mass_coef, double stiff_coef,
                           const uint8_t* node_fixed, skt_operator** out);

/* The two transfers above on a tet4 mesh, taking the node -> element lists
 * and connectivity of a tet4 operator built on it (any physics): volume-
 * weighted element -> node averaging in ascending element order
 * (helmholtz_filter_nodal.py:31-59) and the 4-node mean (:27-28), with the
 * same clamp_nonpositive / mean_cols options. */
int skt_tet4_elem_to_node(const skt_operator* op, int ncols, const double* vol,
                          const uint8_t* design, double fixed_value,
                          const double* rho_elem, double* rho_node, hipStream_t stream);
int skt_tet4_node_to_elem(const skt_operator* op, int ncols, const double* rho_node,
                          int clamp_nonpositive, int mean_cols, double* rho_elem,
                          hipStream_t stream);

/* Robin terms of the heat path on structured hex8 grids
 * (fea/solver_heat.py:552-627,705-980).
 *
 * Nodal densities as objectives_multi_load (:738-745) and
 * _element_to_nodal_average (:600-613) compute them: np.add.at(rho_nodal,
 * t.ravel(), np.repeat(rho, 8)) / count, i.e. node n adds rho[(a n_el + e) /
 * 8] for each of its corner slots a = 0..7 (element e); count (nullable) is
 * the number of elements of the node. */
int skt_hex8_nodal_average_pairing(int64_t nx, int64_t ny, int64_t nz,
                                   const double* rho_elem, double* rho_node,
                                   double* count, hipStream_t stream);

/* get_robin_virtual (:552-573): the 27-point nodal stencil (27, n_nodes) of
 * int h a |grad rho| u v and the load int h a |grad rho| T_env v, a =
 * rho^p (1 - rho)^q, rho the trilinear interpolant of rho_node, Gauss rule
 * of npts (1..4) points per axis. */
int skt_hex8_robin_virtual(int64_t nx, int64_t ny, int64_t nz, double hx,
                           double hy, double hz, int npts,
                           const double* rho_node, double p, double q,
                           double h, double T_env, double* stencil,
                           double* load, hipStream_t stream);

/* Heat-exchange objective terms (fea/solver_heat.py:306-322,380-436,
 * 791-832) on nodal densities rho_node and temperatures T (n_nodes, ncols),
 * npts Gauss points per axis:
 *   den (n_el)          int_e |grad rho_h|                      (J_den parts)
 *   num (n_el, ncols)   int_e (T_c - T_env) |grad rho_h|        (J_num parts
 *                       before the factor -T_env h)
 *   rhs (n_nodes)       int -T_env h_eff |grad rho_h| v, h_eff the
 *                       interpolant of h rho^p (1 - rho)^q (adjoint load
 *                       before the division by J_den). */
int skt_hex8_heat_exchange(int64_t nx, int64_t ny, int64_t nz, double hx,
                           double hy, double hz, int npts,
                           const double* rho_node, int ncols, const double* T,
                           double p, double q, double h, double T_env,
                           double* den, double* num, double* rhs,
                           hipStream_t stream);

/* _robin_compliance_explicit_grad_ (:575-597) per load column of T
 * (n_nodes, ncols), mapped to elements by _nodal_gradient_to_element_gradient
 * (:616-626): out (n_el, ncols); work: n_nodes doubles. */
int skt_hex8_robin_grad(int64_t nx, int64_t ny, int64_t nz, double hx,
                        double hy, double hz, int npts, const double* rho_node,
                        const double* count, int ncols, const double* T,
                        double p, double q, double h, double T_env,
                        double* work, double* out, hipStream_t stream);

/* Attach (stencil != NULL, copied) or drop (NULL) an extra scalar 27-point
 * stencil on a scalar structured hex8 operator: applies add it on free rows
 * and the diagonal includes its centre.  The conduction plus Robin matrix of
 * solve_multi_load (:198-253); the multigrid preconditioner keeps the
 * conduction part. */
int skt_operator_set_extra(skt_operator* op, const double* stencil,
                           hipStream_t stream);

/* Y = S X for a 27-point stencil (27, n_nodes), X, Y (n_nodes, ncols), no
 * Dirichlet rows: the unenforced Robin part of T^T K T (:785-789). */
int skt_stencil27_apply(int64_t nx, int64_t ny, int64_t nz, int ncols,
                        const double* stencil, const double* X, double* Y,
                        hipStream_t stream);

/* Unstructured linear tetrahedra.  coords: (3, n_nodes) float64 device
 * array (x row, y row, z row); conn: (4, n_el) int32 device array.
 * Same operator semantics as above on tet meshes. */
int skt_tet4_create(int physics, int64_t n_nodes, int64_t n_el,
                    const double* coords, const int32_t* conn, double nu,
                    const uint8_t* dof_fixed, skt_operator** out);

/* Declare that a tet4 operator's mesh is a grid of nx*ny*nz cells numbered
 * x-fastest, nodes at the (nx+1)(ny+1)(nz+1) cell corners numbered the same
 * way, and element 6 c + s one of the six tets of cell c (checked).  The
 * solver's multigrid preconditioner then coarsens the grid: its first
 * coarse level is the Galerkin product of the tets' element matrices with
 * trilinear interpolation on the cells (the counterpart of the reference's
 * cg_pyamg option, fea/solver_elastic.py:93-104, for such meshes). */
int skt_tet4_set_grid(skt_operator* op, int64_t nx, int64_t ny, int64_t nz);

int skt_operator_destroy(skt_operator* op);

/* ndof, n_el, ncomp (3 or 1) of the operator. */
int skt_operator_info(const skt_operator* op, int64_t* ndof, int64_t* n_el,
                      int* ncomp);

/* Copy the float64 unit-modulus element matrix of a hex8 operator
 * ((ncomp*8)^2 values, host memory). */
int skt_operator_element_matrix(const skt_operator* op, double* host_out);

/* Y = K_e(E) X for ncols columns.  E, X, Y all of dtype. */
int skt_operator_apply(skt_operator* op, int dtype, int ncols, const void* E,
                       const void* X, void* Y, hipStream_t stream);

/* diag[dof] of K_e(E) (1 on Dirichlet DOFs), dtype. */
int skt_operator_diag(skt_operator* op, int dtype, const void* E, void* diag,
                      hipStream_t stream);

/* energy[e, c] = 1/2 E_e u_e^T K_e^0 u_e for each column of U (float64 out).
 * fea/solver_elastic.py:470-537 (strain_energy_skfem_multi) and
 * fea/solver_heat.py:256-303 (heat_energy_skfem_multi). */
int skt_operator_energy(skt_operator* op, int dtype, int ncols, const void* E,
                        const void* U, double* energy, hipStream_t stream);

/* ------------------------------------------------------------------------
 * Batched preconditioned CG on an operator
 *   Replaces solve_u / solve_multi_load (fea/solver_elastic.py:61-143,
 *   240-395) and solve_heat_system_multi (fea/solver_heat.py:136-195):
 *   all ncols right-hand sides share one operator apply per iteration.
 * mode SKT_F64: float64 CG to ||r|| <= rtol ||F|| per column.
 * mode SKT_F32: float32 PCG with float64 dot products and norms and
 *   reliable updates: once a column's recurrence residual has fallen by
 *   inner_rtol since the last update, the float32 correction is added to U
 *   in float64 and the true residual F - K U (float64 operator) replaces the
 *   recurrence residual; convergence is declared on true residuals only.
 *   inner_rtol <= 0 picks it automatically (1e-2 with multigrid, 1e-1 with
 *   Jacobi). 
 * ---------------------------------------------------------------------- */
int skt_solver_create(skt_operator* op, int max_cols, skt_solver** out);
int skt_solver_destroy(skt_solver* s);

/* E64: (n_el,) float64 moduli; E32: float32 copy (required in SKT_F32 mode,
 * may be NULL in SKT_F64 mode).  F: (ndof, ncols) float64 enforced
 * right-hand side -- F[D] holds the prescribed Dirichlet values (0 for a
 * clamp), as skfem.enforce builds it.  U: (ndof, ncols) float64, the initial
 * guess on entry (warm start; U[D] is overwritten with F[D]) and the
 * solution on exit.
 * iters[0] = PCG iterations, iters[1] = reliable updates (0 in SKT_F64);
 * relres[c] = final ||F - K U|| / ||F|| per column (host arrays, valid on
 * return).  The copy of the solution into U is stream-ordered on `stream`
 * and may still be running when the call returns. */
int skt_solver_solve(skt_solver* s, int mode, int ncols, const double* E64,
                     const float* E32, const double* F, double* U,
                     double rtol, int maxiter, double inner_rtol,
                     int64_t* iters, double* relres, hipStream_t stream);

/* Preconditioner selection (the reference's iterative options are scipy
 * cg with Jacobi or pyamg, fea/solver_elastic.py:84-101):
 *   SKT_PC_JACOBI  diagonal;
 *   SKT_PC_MG      geometric multigrid V-cycle on structured hex8 grids:
 *                  Galerkin coarse operators of trilinear interpolation with
 *                  the Dirichlet DOFs removed, Chebyshev-Jacobi smoothing of
 *                  smooth_degree steps, dense inverse on the coarsest level
 *                  (at most max_coarse DOFs);
 *   SKT_PC_AUTO    multigrid where the grid allows it, else Jacobi (default).
 * skt_solver_info reports the preconditioner the last solve used and its
 * number of levels (1 for Jacobi). */
#define SKT_PC_JACOBI 0
#define SKT_PC_MG 1
#define SKT_PC_AUTO 2
int skt_solver_configure(skt_solver* s, int precond, int smooth_degree,
                         int max_coarse);
int skt_solver_info(const skt_solver* s, int* precond_used, int* levels);

/* Measurement hooks (no reference counterpart; used by the benchmark):
 * with timing enabled, every operator apply inside the CG loop is bracketed
 * by hipEvents on the solve stream.  get_timing returns the accumulated
 * milliseconds and launch counts per dtype ([0] float64, [1] float32). */
int skt_solver_set_timing(skt_solver* s, int enable);
int skt_solver_get_timing(skt_solver* s, double* apply_ms, int64_t* n_apply,
                          int reset);
/* With timing enabled, the multigrid V-cycle's fine-level launches on a
 * structured hex8 grid are bracketed too: ms / n ([4] each) per variant --
 * [0] the fused first two Chebyshev steps, [1] the residual, [2] a
 * Chebyshev step, [3] the last Chebyshev step with its fused r . z. */
int skt_solver_get_vcycle_timing(skt_solver* s, double* ms, int64_t* n, int reset);

/* The PCG iterations between two convergence polls, and the multigrid
 * set-up once it is warm, run as captured hipGraphs replayed on the solve
 * stream (no reference counterpart: the reference's solves are host library
 * calls).  graph_stats returns the replays and captures so far.  SKT_GRAPH=0 in the environment when the
 * solver is created makes it launch the same kernels one by one instead;
 * timing (above) also does. */
int skt_solver_graph_stats(const skt_solver* s, int64_t* launches, int64_t* captures);

/* ------------------------------------------------------------------------
 * Gaussian density filter on element centres (filters/spacial.py)
 * ---------------------------------------------------------------------- */

/* Build the row-normalised neighbour-list (CSR) filter over design
 * elements: neighbours j of design element i are design elements with
 * dx^2+dy^2+dz^2 <= support^2 (cKDTree.query_ball_point rule), weights
 * exp(-0.5 (d/sigma)^2), columns in the compressed design numbering and
 * ascending order, rows scaled by 1/rowsum (spacial.py:51-103).
 * centers: (3, n_el) float64 device array; design: (n_el,) uint8. */
int skt_filter_create(int64_t n_el, const double* centers,
                      const uint8_t* design, double sigma, double support,
                      skt_filter** out);
int skt_filter_destroy(skt_filter* f);
int skt_filter_info(const skt_filter* f, int64_t* n_design, int64_t* nnz);

/* Copy the CSR to device buffers: indptr (n_design+1) int64,
 * indices (nnz) int32, values (nnz) float64, design_ids (n_design) int32. */
int skt_filter_export(const skt_filter* f, int64_t* indptr, int32_t* indices,
                      double* values, int32_t* design_ids, hipStream_t stream);

/* The filter of a structured nx x ny x nz grid of hx x hy x hz boxes
 * (elements x-fastest) in stencil form: offsets inside the support with
 * one Gaussian weight each, times the row's 1 / row-sum.  The form is only
 * switched on after every CSR entry has been checked equal to it (pattern
 * and weight bits), and forward / gradient / skt_oc_bisect then read no
 * column indices or per-entry weights; *active reports whether it is on.
 * Same matrix as spacial.py:51-103 builds; an optimisation of this build. */
int skt_filter_set_grid(skt_filter* f, int64_t nx, int64_t ny, int64_t nz,
                        double hx, double hy, double hz, int* active);

/* spacial.py:135-157: y = x; y[design] = W x[design]  (n_el arrays). */
int skt_filter_forward(skt_filter* f, const double* x, double* y,
                       hipStream_t stream);

/* spacial.py:159-166: g = 0; g[design] = W^T v[design]. */
int skt_filter_gradient(skt_filter* f, const double* v, double* g,
                        hipStream_t stream);

/* ------------------------------------------------------------------------
 * OC update: lambda bisection on the physical volume, all on the device
 *   core/optimizers/oc.py:16-95 (bisection_with_physical_volume), driven as
 *   in oc.py:199-228.  Every step: candidate update, filter gather,
 *   Heaviside projection and float64 volume sum; the bracket update and
 *   exit test run on the device (no host round trip).
 * rho_full (n_el): current densities (the filter couples design elements
 *   only, so the volume depends on the design candidates alone and the
 *   non-design entries are not read);
 * rho_e, dC, vol: (n_design,) in design order; rho_out (n_design,): the
 * candidate at the final lambda.  result (device, 4 doubles):
 * {lmid, vol_error, n_evals, l2 - l1}.
 * ---------------------------------------------------------------------- */
int skt_oc_bisect(skt_filter* f, const double* rho_full, const double* rho_e,
                  const double* dC, const double* vol, double vol_sum,
                  double vol_frac, double move_limit, double eta,
                  double rho_min, double rho_max, double rate_min,
                  double rate_max, double beta, double beta_eta, double l1,
                  double l2, double tolerance, double vol_tol, int max_iter,
                  double* rho_out, double* result, hipStream_t stream);

/* core/optimizers/oc.py:187-196: s = max(*scale, floor); running_scale =
 * first ? s : keep*running_scale + (1-keep)*s; out = dC / running_scale.
 * scale and running_scale are device scalars (no host round trip). */
int skt_oc_normalize(int64_t n, const double* dC, const double* scale,
                     double floor_, double keep, int first, double* running_scale,
                     double* out, hipStream_t stream);

/* core/optimizers/oc.py:230-240: KKT residual max |dC_raw + lmid*rs*dV| over
 * rho in (rho_min+1e-6, rho_max-1e-6); lmid = bisect_result[0],
 * rs = running_scale[0]; 0 when no element qualifies.  out: device scalar. */
int skt_oc_kkt(int64_t n, const double* dC_raw, const double* rho, const double* dV,
               const double* bisect_result, const double* running_scale,
               double rho_min, double rho_max, double* out, hipStream_t stream);

/* logmoc.py:36-66 lagrangian_log_update, in place on rho (n,). */
int skt_logmoc_update(int64_t n, double* rho, const double* dL, double eta,
                      double move_limit, double rho_min, double rho_max,
                      double clip, hipStream_t stream);

/* Gather / scatter between full (n_el) and design-ordered (n_design) arrays. */
int skt_filter_gather_design(skt_filter* f, const double* full, double* design,
                             hipStream_t stream);
int skt_filter_scatter_design(skt_filter* f, const double* design, double* full,
                              hipStream_t stream);

/* The same with an explicit list of design element ids (int32, ascending):
 * design[r] = full[ids[r]] and full[ids[r]] = design[r]
 * (common_density.py:1108-1109, oc.py:67-68). */
int skt_design_gather(int64_t n_design, const int32_t* design_ids,
                      const double* full, double* design, hipStream_t stream);
int skt_design_scatter(int64_t n_design, const int32_t* design_ids,
                       const double* design, double* full, hipStream_t stream);

/* One OC candidate at multiplier lmid (oc.py:51-68): rho_out (n_design) =
 * clip(rho_e * clip((-dC / (lmid + 1e-12))^eta, rate_min, rate_max),
 * max(rho_e - move, rho_min), min(rho_e + move, rho_max)), also written to
 * rho_full[ids] when rho_full is not NULL.  The building block of the
 * bisection for filters that are not a sparse matrix (the host drives the
 * bracket and runs the filter between candidates). */
int skt_oc_candidate(int64_t n_design, const int32_t* design_ids,
                     const double* rho_e, const double* dC, double lmid,
                     double eta, double rho_min, double rho_max,
                     double rate_min, double rate_max, double move_limit,
                     double* rho_full, double* rho_out, hipStream_t stream);

/* Robin terms of the heat path on tet4 meshes (fea/solver_heat.py:306-624,
 * 705-980) with P1 fields, for a scalar tet4 operator (its node -> element
 * lists, gradients and volumes).  order selects the tetrahedral rule of the
 * volume forms: 1 the centroid, 2 the symmetric 4-point rule of degree 2
 * (the rule of a P1 basis' default integration order 2).
 *
 * Nodal densities of objectives_multi_load (:738-745): np.add.at(rho_nodal,
 * t.ravel(), np.repeat(rho, 4)) / count, i.e. node n adds rho[(a n_el + e) /
 * 4] for its corner slots a = 0..3 in that order; count (nullable). */
int skt_tet4_nodal_average_pairing(const skt_operator* op, const double* rho_elem,
                                   double* rho_node, double* count, hipStream_t stream);

/* get_robin_virtual (:552-573): elem_mat (10, n_el), the element matrices of
 * int h a |grad rho| u v packed 00 11 22 33 01 02 03 12 13 23, and load
 * (n_nodes) of int h a |grad rho| T_env v, a = rho^p (1 - rho)^q; work:
 * 4 n_el doubles. */
int skt_tet4_robin_virtual(const skt_operator* op, int order, const double* rho_node,
                           double p, double q, double h, double T_env, double* elem_mat,
                           double* load, double* work, hipStream_t stream);

/* Heat-exchange terms (:306-322,380-436,791-832) as skt_hex8_heat_exchange,
 * exact for P1 fields; work: 4 n_el doubles. */
int skt_tet4_heat_exchange(const skt_operator* op, const double* rho_node, int ncols,
                           const double* T, double p, double q, double h, double T_env,
                           double* den, double* num, double* rhs, double* work,
                           hipStream_t stream);

/* _robin_compliance_explicit_grad_ (:575-597) per load column of T
 * (n_nodes, ncols), mapped to elements by _nodal_gradient_to_element_gradient
 * (:616-626): out (n_el, ncols); work: 4 n_el + n_nodes doubles. */
int skt_tet4_robin_grad(const skt_operator* op, int order, const double* rho_node,
                        const double* count, int ncols, const double* T, double p, double q,
                        double h, double T_env, double* work, double* out, hipStream_t stream);

/* Attach (elem_mat != NULL, copied; (10, n_el) packed as above) or drop
 * (NULL) extra element matrices on a scalar tet4 operator: applies add R_e x_e
 * on free rows and the diagonal includes R_aa -- the conduction plus Robin
 * matrix of solve_multi_load (:198-253); a multigrid hierarchy keeps the
 * conduction part. */
int skt_tet4_set_extra(skt_operator* op, const double* elem_mat, hipStream_t stream);

/* out (n_el, ncols) = x_e^T R_e x_e of the attached extra element matrices,
 * X (n_nodes, ncols): the Robin part of T^T K T (:785-789). */
int skt_tet4_extra_quadform(const skt_operator* op, int ncols, const double* X, double* out,
                            hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SKTOPT_AMD_H */
