// Geometric multigrid preconditioner for structured hex8 operators (internal).
//
// Level 0 is the matrix-free fine operator.  Each coarser level groups
// fx*fy*fz fine elements into one coarse element (factors 2, 3 or 5 per axis,
// 1 where an axis cannot be divided) and carries the exact Galerkin operator
// P^T A P of trilinear interpolation, built element by element: the coarse
// element matrix is sum_children P_c^T K_c P_c, with the fine Dirichlet rows
// and columns removed first, so the hierarchy preconditions the operator on
// the free DOFs.  Coarse levels store the assembled 27-point node stencil
// (structure of arrays, [27*NC*NC][nodes]); the coarsest level is inverted
// densely.  Smoothing is Chebyshev-accelerated Jacobi with the spectral bound
// of D^-1 A estimated on the device by power iteration, so a V-cycle is a
// fixed symmetric linear operator and never waits on the host.  Set-ups for
// new moduli restart the power iteration from the previous set-up's vector
// (the moduli of successive optimisation steps differ little), with fewer
// steps.
#pragma once

#include <vector>

#include "operator.hpp"

namespace skt {

struct MGLevel {
  int64_t nx = 0, ny = 0, nz = 0;     // elements
  int64_t nn = 0, ndof = 0;           // nodes, DOFs
  int f[3] = {1, 1, 1};               // factor from the finer level
  double* Kel = nullptr;              // (n_el, ND*ND) Galerkin element matrices, fp64
  void* st = nullptr;                 // stencil in T (levels 1..L-1), node- or tile-major
  int layout = 0;                     // stencil layout (node-major, big or small tiles)
  void* dinv = nullptr;               // (ndof) in T
  void* b = nullptr;                  // right-hand side (ndof*k, blocked)
  void* x = nullptr;                  // iterate
  void* w = nullptr;                  // operator output
  void* d = nullptr;                  // Chebyshev direction
  void* rt = nullptr;                 // restriction staging (x-y restricted fine planes)
};

// Levels with at least this many nodes use the big stencil tiles, smaller
// ones the small tiles.
constexpr int64_t kTiledNodes = 20000;

struct MGData {
  int dtype = -1;                     // SKT_F64 / SKT_F32 the buffers hold
  int max_cols = 0;
  std::vector<MGLevel> lv;            // lv[0]: fine level (vector: x only)
  double* Ad = nullptr;               // coarsest dense matrix, fp64 (n x n), ping
  double* Ad2 = nullptr;              // inverses of the current and next pivot blocks
  void* Ainv = nullptr;               // coarsest inverse in T
  double* lam = nullptr;              // per level lambda_max(D^-1 A) estimates
  double* coef = nullptr;             // per level Chebyshev (c1, c2) per step
  double* part = nullptr;             // power-iteration partials
  double* K0b = nullptr;              // unit element matrix, corner-bit order
  double* Mch = nullptr;              // P_j^T K0 P_j per child position of level 1
  uint8_t* gflag = nullptr;           // level-1 elements with Dirichlet DOFs below
  hipStream_t side = nullptr;         // set-up: Galerkin products, coarsest inverse
  hipStream_t side2 = nullptr;        // set-up: stencils and Jacobi inverses
  hipEvent_t ev_fork = nullptr, ev_gal = nullptr, ev_st = nullptr, ev_join = nullptr;
  void* pv = nullptr;                 // power-iteration scratch (2 * ndof0) in T
  void* pvw = nullptr;                // per level (levels 0..L-1, concatenated) the last
                                      // normalised power-iteration vector in T
  int warm = 0;                       // pvw holds vectors from an earlier set-up
  int degree = 2;
  double ratio = 10.0;
  int power_iters = 20;               // steps from the hashed start vector
  int power_iters_warm = 5;           // steps from the previous set-up's vector
  uint64_t gen = 0;                   // bumped whenever the buffers are (re)allocated
  // Lagged coarsest inverse: once an inverse exists, a set-up only assembles
  // the coarsest matrix; the solve that follows uses the inverse of the
  // previous set-up's matrix, and the new matrix is inverted on `gjs` during
  // that solve (mg_lag_post), to be cast into Ainv by the next set-up
  // (mg_lag_pre).  On with SKT_MG_LAG=1.
  int lag = 0;
  int have_inv = 0;                   // Ainv holds an inverse
  // A set-up may keep the previous coarsest inverse (reuse_inv, decided per
  // set-up by the solver); it stays an SPD coarse solve, so the V-cycle is a
  // fixed symmetric operator all the same.  An inverse serves at most
  // `coarse_refresh` set-ups (1: a fresh one every time;
  // SKT_MG_COARSE_REFRESH=n overrides).
  int coarse_refresh = 2;
  int reuse_inv = 0;
  int inv_pending = 0;                // Ad holds, or is becoming, a newer inverse
  hipStream_t gjs = nullptr;
  hipEvent_t ev_asm = nullptr, ev_gj = nullptr;
  // Optional HIP-event timing of the fine level's fused modal launches
  // (the solver's timing mode, direct launches): event pairs tagged with
  // the variant (MG_T_*), folded into per-variant totals by mg_collect_timing.
  static constexpr int kTimedMax = 256;
  hipEvent_t tev[2 * kTimedMax] = {};
  int ttag[kTimedMax] = {};
  int tcount = 0;
  int time_fine = 0;
  double tms[4] = {0, 0, 0, 0};
  int64_t tn[4] = {0, 0, 0, 0};
};

// fine-level modal variants timed by the V-cycle
enum { MG_T_CHEB0 = 0, MG_T_RESID = 1, MG_T_CHEB = 2, MG_T_CHEB_DOT = 3 };

struct MGConfig {
  int degree = 2;          // Chebyshev steps per smoothing
  int max_coarse = 2000;   // coarsest level at most this many DOFs
  double ratio = 10.0;     // smoothing interval [lambda/ratio, lambda]
  int power_iters = 20;
};

// Build the level structure for an operator (no numeric data yet).  Returns
// false when the grid cannot be coarsened to at most max_coarse DOFs.
bool mg_plan(const skt_operator* op, const MGConfig& cfg, std::vector<MGLevel>* out);

// Allocate buffers for dtype and k columns (idempotent for the same dtype).
int mg_alloc(MGData* mg, const skt_operator* op, const MGConfig& cfg, int dtype, int max_cols);
void mg_free(MGData* mg);

// Numeric setup for the moduli E (fp64) -- Galerkin hierarchy, stencils,
// coarsest inverse, smoother bounds.  dinv0 is the fine Jacobi inverse in T.
// Power-iteration steps of a warm set-up (SKT_MG_POWER_WARM overrides).
int mg_warm_steps(const MGData* mg);
// Set-ups one coarsest inverse may serve (SKT_MG_COARSE_REFRESH overrides).
int mg_coarse_refresh(const MGData* mg);
int mg_setup(MGData* mg, const skt_operator* op, const double* E64, const void* E, const void* dinv0,
             hipStream_t s);

// Lagged coarsest inverse (see MGData::lag).  mg_lagged: the next set-up
// assembles the coarsest matrix only.  mg_lag_pre, before that set-up and
// outside any stream capture: `s` waits for the background inverse and casts
// it into Ainv.  mg_lag_post, after it: the assembled matrix is inverted on
// the multigrid's own stream once `s` reaches this point.
bool mg_lagged(const MGData* mg);

// Timing of the fine-level V-cycle launches (creates the events on first
// use); mg_collect_timing folds completed event pairs into tms / tn.
int mg_set_timing(MGData* mg, int enable);
int mg_collect_timing(MGData* mg);
int mg_lag_pre(MGData* mg, hipStream_t s);
int mg_lag_post(MGData* mg, hipStream_t s);

// z = V(r) for ncols columns in the blocked layout [ncols/kc][ndof][kc];
// w0 is fine-level scratch of the same size.  With dot, the last smoothing
// step also leaves the per-column partials of r . z there when they fit in
// max_slots per column (*slots = their count, else 0).
// With stop (column stop flags, kStopStride apart), the fine-level and
// tiled coarse-level kernels may skip workgroups whose columns have all
// stopped; z is then stale for those columns.
int mg_vcycle(MGData* mg, const skt_operator* op, const void* E, const void* dinv0, int ncols,
              int kc, const void* r, void* z, void* w0, hipStream_t s, double* dot = nullptr,
              int max_slots = 0, int* slots = nullptr, const double* stop = nullptr);

}  // namespace skt
