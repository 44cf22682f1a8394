// Gaussian density filter as a precomputed neighbour-list (CSR) gather, and
// the device-side OC lambda bisection that runs the filter once per step.
//
// Setup (skt_filter_create) bins element centres on a uniform grid, finds
// neighbours with the inclusion rule dx^2+dy^2+dz^2 <= support^2 evaluated
// in IEEE float64 without contraction, sorts every row by column, and stores
//   row_ptr (int64), gcol (int32 full element index), W and W^T values.
// The neighbour relation is symmetric, so W^T has W's pattern; its values
// are placed by a binary search of the mirrored entry.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "common.hpp"

struct skt_filter {
  int64_t n_el = 0;
  int64_t n_design = 0;
  int64_t nnz = 0;
  int32_t* design_ids = nullptr;   // (n_design)
  int32_t* colmap = nullptr;       // (n_el) compressed index or -1
  int64_t* row_ptr = nullptr;      // (n_design + 1)
  int32_t* gcol = nullptr;         // (nnz) full element index of the column
  double* w = nullptr;             // (nnz) row-normalised weights
  double* wt = nullptr;            // (nnz) transposed weights, same pattern
  double* inv_rs = nullptr;        // (n_design) 1 / row sum of the Gaussian weights
  // structured-grid form (skt_filter_set_grid): the same matrix as a
  // stencil of element offsets with per-offset weights, used when it has
  // been verified equal to the CSR entry by entry
  int stencil = 0;
  double sigma = 0.0, support = 0.0;
  int64_t gnx = 0, gny = 0, gnz = 0;
  int n_off = 0;
  int R = 0;                       // stencil half-width in elements
  int32_t* off = nullptr;          // (n_off) packed dx, dy, dz, ascending element order
  int32_t* del = nullptr;          // (n_off) element-index delta
  double* u_off = nullptr;         // (n_off) unnormalised Gaussian weight
  // bisection workspace
  double* cand_full = nullptr;     // (n_el)
  double* cand_tree = nullptr;     // (n_el, kTreeStride): candidates at the tree's midpoints
  double* part_tree = nullptr;     // (kTreeK, kPartMax)
  double* part = nullptr;          // (kPartMax)
  double* state = nullptr;         // bisection state
};

namespace skt {

constexpr int kPartMax = 4096;
// CSR bisection: each pass evaluates the midpoints of the next kTreeLevels
// levels of the bisection tree (kTreeK of them) in one read of the matrix
constexpr int kTreeLevels = 3;
constexpr int kTreeK = (1 << kTreeLevels) - 1;
constexpr int kTreeStride = 8;
static_assert(kTreeK <= kTreeStride, "tree candidates per element");
constexpr int kSeg = 8;  // lanes per CSR row in the gathers

__device__ __forceinline__ double dist2(const double* __restrict__ c, int64_t n,
                                        int64_t i, int64_t j) {
#pragma clang fp contract(off)
  double dx = c[j] - c[i];
  double dy = c[n + j] - c[n + i];
  double dz = c[2 * n + j] - c[2 * n + i];
  double s = dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return s;
}

__global__ void k_cell_of(int64_t n, const double* __restrict__ c, double x0,
                          double y0, double z0, double inv, int64_t gx,
                          int64_t gy, int64_t gz, int64_t* __restrict__ cell,
                          int32_t* __restrict__ ids) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t a = (int64_t)floor((c[i] - x0) * inv);
    int64_t b = (int64_t)floor((c[n + i] - y0) * inv);
    int64_t d = (int64_t)floor((c[2 * n + i] - z0) * inv);
    a = a < 0 ? 0 : (a >= gx ? gx - 1 : a);
    b = b < 0 ? 0 : (b >= gy ? gy - 1 : b);
    d = d < 0 ? 0 : (d >= gz ? gz - 1 : d);
    cell[i] = a + gx * (b + gy * d);
    ids[i] = (int32_t)i;
  }
}

__global__ void k_cell_bounds(int64_t n, const int64_t* __restrict__ cell_sorted,
                              int64_t* __restrict__ start, int64_t* __restrict__ end) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t c = cell_sorted[i];
    if (i == 0 || cell_sorted[i - 1] != c) start[c] = i;
    if (i == n - 1 || cell_sorted[i + 1] != c) end[c] = i + 1;
  }
}

__global__ void k_design_flags(int64_t n, const uint8_t* __restrict__ design,
                               int32_t* __restrict__ flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = design[i] ? 1 : 0;
}

__global__ void k_design_ids(int64_t n, const uint8_t* __restrict__ design,
                             const int32_t* __restrict__ pos,
                             int32_t* __restrict__ ids, int32_t* __restrict__ colmap) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (design[i]) {
      ids[pos[i]] = (int32_t)i;
      colmap[i] = pos[i];
    } else {
      colmap[i] = -1;
    }
  }
}

// count (fill == false) or fill (fill == true) the neighbours of each row
template <bool kFill>
__global__ void k_neighbours(int64_t n_el, int64_t n_design,
                             const double* __restrict__ c,
                             const int32_t* __restrict__ design_ids,
                             const int32_t* __restrict__ colmap,
                             const int64_t* __restrict__ cell,
                             const int32_t* __restrict__ sorted_ids,
                             const int64_t* __restrict__ cstart,
                             const int64_t* __restrict__ cend, int64_t gx,
                             int64_t gy, int64_t gz, double r2,
                             int64_t* __restrict__ counts,
                             const int64_t* __restrict__ row_ptr,
                             int32_t* __restrict__ gcol) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_design;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = design_ids[r];
    const int64_t ci = cell[i];
    const int64_t a = ci % gx, b = (ci / gx) % gy, d = ci / (gx * gy);
    int64_t cnt = 0;
    int64_t out = kFill ? row_ptr[r] : 0;
    for (int64_t dd = d - 1; dd <= d + 1; ++dd) {
      if (dd < 0 || dd >= gz) continue;
      for (int64_t bb = b - 1; bb <= b + 1; ++bb) {
        if (bb < 0 || bb >= gy) continue;
        for (int64_t aa = a - 1; aa <= a + 1; ++aa) {
          if (aa < 0 || aa >= gx) continue;
          const int64_t cc = aa + gx * (bb + gy * dd);
          const int64_t s0 = cstart[cc], s1 = cend[cc];
          for (int64_t q = s0; q < s1; ++q) {
            const int64_t j = sorted_ids[q];
            if (colmap[j] < 0) continue;
            if (dist2(c, n_el, i, j) <= r2) {
              if (kFill) gcol[out++] = (int32_t)j;
              ++cnt;
            }
          }
        }
      }
    }
    if (!kFill) counts[r] = cnt;
  }
}

// Row-wise insertion sort of gcol (rows are short), then Gaussian weights and
// row normalisation in ascending column order.
__global__ void k_sort_weights(int64_t n_el, int64_t n_design,
                               const double* __restrict__ c,
                               const int32_t* __restrict__ design_ids,
                               const int64_t* __restrict__ row_ptr,
                               int32_t* __restrict__ gcol, double sigma,
                               double* __restrict__ w, double* __restrict__ inv_rs) {
#pragma clang fp contract(off)
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_design;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s0 = row_ptr[r], s1 = row_ptr[r + 1];
    for (int64_t q = s0 + 1; q < s1; ++q) {
      int32_t v = gcol[q];
      int64_t p = q - 1;
      while (p >= s0 && gcol[p] > v) {
        gcol[p + 1] = gcol[p];
        --p;
      }
      gcol[p + 1] = v;
    }
    const int64_t i = design_ids[r];
    double sum = 0.0;
    for (int64_t q = s0; q < s1; ++q) {
      double d = sqrt(dist2(c, n_el, i, gcol[q]));
      double t = d / sigma;
      double wq = exp(-0.5 * (t * t));
      w[q] = wq;
      sum = sum + wq;
    }
    if (sum == 0.0) sum = 1.0;
    const double inv = 1.0 / sum;
    inv_rs[r] = inv;
    for (int64_t q = s0; q < s1; ++q) w[q] = inv * w[q];
  }
}

__global__ void k_transpose_values(int64_t n_design, const int32_t* __restrict__ design_ids,
                                   const int32_t* __restrict__ colmap,
                                   const int64_t* __restrict__ row_ptr,
                                   const int32_t* __restrict__ gcol,
                                   const double* __restrict__ w,
                                   double* __restrict__ wt) {
  // For entry (r -> j) at q, W^T[j, r] = W[r, j]; W^T row j has W row j's
  // pattern, so locate column design_ids[r] in row colmap[j].
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_design;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t gi = design_ids[r];
    for (int64_t q = row_ptr[r]; q < row_ptr[r + 1]; ++q) {
      const int64_t rj = colmap[gcol[q]];
      int64_t lo = row_ptr[rj], hi = row_ptr[rj + 1] - 1;
      while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (gcol[mid] < gi) lo = mid + 1; else hi = mid;
      }
      wt[lo] = w[q];
    }
  }
}

// y[e] = x[e] for non-design e; y[design row] = sum_q val[q] x[gcol[q]].
// kSeg lanes share one row; fixed butterfly reduction inside the segment.
__global__ void k_csr_apply(int64_t n_el, const int32_t* __restrict__ colmap,
                            const int64_t* __restrict__ row_ptr,
                            const int32_t* __restrict__ gcol,
                            const double* __restrict__ val,
                            const double* __restrict__ x,
                            double* __restrict__ y, int zero_nondesign) {
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int sub = threadIdx.x & (kSeg - 1);
  for (int64_t e = tid / kSeg; e < n_el; e += ((int64_t)gridDim.x * blockDim.x) / kSeg) {
    const int32_t r = colmap[e];
    double acc = 0.0;
    if (r >= 0) {
      const int64_t s0 = row_ptr[r], s1 = row_ptr[r + 1];
      for (int64_t q = s0 + sub; q < s1; q += kSeg) acc += val[q] * x[gcol[q]];
    }
#pragma unroll
    for (int off = kSeg / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kSeg);
    if (sub == 0) y[e] = (r >= 0) ? acc : (zero_nondesign ? 0.0 : x[e]);
  }
}

__global__ void k_gather_design(int64_t n_design, const int32_t* __restrict__ ids,
                                const double* __restrict__ full,
                                double* __restrict__ des) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_design;
       r += (int64_t)gridDim.x * blockDim.x)
    des[r] = full[ids[r]];
}

__global__ void k_scatter_design(int64_t n_design, const int32_t* __restrict__ ids,
                                 const double* __restrict__ des,
                                 double* __restrict__ full) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_design;
       r += (int64_t)gridDim.x * blockDim.x)
    full[ids[r]] = des[r];
}

// ------------------------------------------------------- structured stencil
// On a structured nx x ny x nz grid the filter is a stencil: element e's row
// holds the design elements e + o for the offsets o of the Gaussian support,
// with weight inv_rs[row] * u[o].  The weights below are the same
// floating-point operations as k_sort_weights applied to the offset's centre
// difference, and skt_filter_set_grid only switches to this form after
// checking every CSR entry against it.

// u[o] and the inclusion flag for every offset of a (2R+1)^3 box
__global__ void k_offset_weights(int R, double hx, double hy, double hz, double r2,
                                 double sigma, double* __restrict__ u,
                                 int32_t* __restrict__ inside) {
#pragma clang fp contract(off)
  const int w = 2 * R + 1;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < w * w * w;
       o += gridDim.x * blockDim.x) {
    const double dx = (double)(o % w - R) * hx, dy = (double)((o / w) % w - R) * hy,
                 dz = (double)(o / (w * w) - R) * hz;
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    inside[o] = s <= r2 ? 1 : 0;
    double d = sqrt(s);
    double t = d / sigma;
    u[o] = exp(-0.5 * (t * t));
  }
}

// 1 where a CSR row differs from the stencil (pattern or weight bits)
__global__ void k_check_stencil(int64_t n_design, int64_t nx, int64_t ny, int64_t nz, int R,
                                const int32_t* __restrict__ design_ids,
                                const int32_t* __restrict__ colmap,
                                const int64_t* __restrict__ row_ptr,
                                const int32_t* __restrict__ gcol,
                                const double* __restrict__ w,
                                const double* __restrict__ inv_rs,
                                const double* __restrict__ ubox,
                                const int32_t* __restrict__ inside,
                                int32_t* __restrict__ bad) {
  const int wd = 2 * R + 1;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_design;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = design_ids[r];
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    int ok = 1;
    // every CSR entry is a stencil offset with the stencil's weight
    for (int64_t q = row_ptr[r]; q < row_ptr[r + 1]; ++q) {
      const int64_t j = gcol[q];
      const int64_t dx = j % nx - ex, dy = (j / nx) % ny - ey, dz = j / (nx * ny) - ez;
      if (dx < -R || dx > R || dy < -R || dy > R || dz < -R || dz > R) { ok = 0; break; }
      const int o = (int)((dx + R) + wd * ((dy + R) + wd * (dz + R)));
      if (!inside[o] || inv_rs[r] * ubox[o] != w[q]) { ok = 0; break; }
    }
    // and the stencil has no design neighbour the CSR lacks
    int64_t cnt = 0;
    for (int dz = -R; dz <= R; ++dz)
      for (int dy = -R; dy <= R; ++dy)
        for (int dx = -R; dx <= R; ++dx) {
          const int o = (dx + R) + wd * ((dy + R) + wd * (dz + R));
          if (!inside[o]) continue;
          const int64_t x2 = ex + dx, y2 = ey + dy, z2 = ez + dz;
          if (x2 < 0 || x2 >= nx || y2 < 0 || y2 >= ny || z2 < 0 || z2 >= nz) continue;
          if (colmap[x2 + nx * (y2 + ny * z2)] >= 0) ++cnt;
        }
    if (cnt != row_ptr[r + 1] - row_ptr[r]) ok = 0;
    if (!ok) bad[0] = 1;
  }
}

// Stencil row sum for element e (design row r): sum_o inv_rs[r_row] u[o] x[e+o]
// over in-grid design neighbours, in ascending element order.  With
// transpose, the row index of the weight is the neighbour's (W^T).  off[o]
// packs (dx + 64) | (dy + 64) << 8 | (dz + 64) << 16 and del[o] is the
// element-index delta.  Indices fit in 32 bits (n_el < 2^31).  kDesignX: x
// is zero on every non-design element, so the design test of the neighbour
// (a second gather) is left out; the terms it would drop are then exact
// zeros and the sum is the same.
template <bool kTranspose, bool kDesignX = false>
__device__ __forceinline__ double stencil_row(int e, int32_t r, int nx, int ny, int nz, int R,
                                              int n_off, const int32_t* __restrict__ off,
                                              const int32_t* __restrict__ del,
                                              const double* __restrict__ u,
                                              const int32_t* __restrict__ colmap,
                                              const double* __restrict__ inv_rs,
                                              const double* __restrict__ x) {
  const int nxy = nx * ny;
  const int ex = e % nx, ey = (e / nx) % ny, ez = e / nxy;
  const double ir = kTranspose ? 0.0 : inv_rs[r];
  double acc = 0.0;
  // branch-free over all offsets (lanes near a face take the same path as
  // interior ones): an out-of-grid neighbour reads the element itself and a
  // selected zero, so loads never wait on a branch and several offsets are
  // in flight at once (four per unrolled step; twelve, three waits instead
  // of nine over the 33 offsets of r = 2, ran slower: 17.5 against 14.7 us
  // per bisection evaluation at 120x60x60)
#pragma unroll 4
  for (int o = 0; o < n_off; ++o) {
    const int pk = off[o];
    const int x2 = ex + (pk & 255) - 64, y2 = ey + ((pk >> 8) & 255) - 64,
              z2 = ez + ((pk >> 16) & 255) - 64;
    const bool inb = (unsigned)x2 < (unsigned)nx && (unsigned)y2 < (unsigned)ny &&
                     (unsigned)z2 < (unsigned)nz;
    const int j = inb ? e + del[o] : e;
    const double xv = x[j];
    if constexpr (kDesignX && !kTranspose) {
      const double term = (ir * u[o]) * xv;
      acc += inb ? term : 0.0;
    } else {
      const int32_t rj = colmap[j];
      double wgt;
      if constexpr (kTranspose)
        wgt = inv_rs[rj < 0 ? 0 : rj] * u[o];
      else
        wgt = ir * u[o];
      const double term = wgt * xv;
      acc += (inb && rj >= 0) ? term : 0.0;
    }
  }
  return acc;
}

template <bool kTranspose>
__global__ void k_stencil_filter(int n_el, int nx, int ny, int nz, int R, int n_off,
                                 const int32_t* __restrict__ off,
                                 const int32_t* __restrict__ del, const double* __restrict__ u,
                                 const int32_t* __restrict__ colmap,
                                 const double* __restrict__ inv_rs,
                                 const double* __restrict__ x, double* __restrict__ y) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n_el; e += gridDim.x * blockDim.x) {
    const int32_t r = colmap[e];
    if (r < 0) {
      y[e] = kTranspose ? 0.0 : x[e];
      continue;
    }
    y[e] = stencil_row<kTranspose>(e, r, nx, ny, nz, R, n_off, off, del, u, colmap, inv_rs, x);
  }
}

// ------------------------------------------------------------------ bisection
// state: [0] l1 [1] l2 [2] lmid [3] vol_error [4] iter [5] done [6] n_evals

__global__ void k_fill(int64_t n, double v, double* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = v;
}

__global__ void k_oc_init(double* st, double l1, double l2) {
  st[0] = l1;
  st[1] = l2;
  st[2] = 0.5 * (l1 + l2);
  st[3] = 0.0;
  st[4] = 0.0;
  st[5] = 0.0;
  st[6] = 0.0;
}


__global__ void k_oc_candidate(int64_t n_design, const int32_t* __restrict__ ids,
                               const double* __restrict__ rho_e,
                               const double* __restrict__ dC,
                               const double* __restrict__ st, double eps,
                               double eta, double rate_min, double rate_max,
                               double move_limit, double rho_min, double rho_max,
                               double* __restrict__ cand_full,
                               double* __restrict__ rho_out) {
  if (st[5] != 0.0) return;
  const double lmid = st[2];
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_design;
       r += (int64_t)gridDim.x * blockDim.x) {
    const double c = oc_candidate(dC[r], rho_e[r], lmid, eps, eta, rate_min, rate_max,
                                  move_limit, rho_min, rho_max);
    rho_out[r] = c;
    cand_full[ids[r]] = c;
  }
}

// Filter gather on the candidate, projection, volume partial sums.
__global__ void k_oc_volume(int64_t n_design, const int32_t* __restrict__ ids,
                            const int64_t* __restrict__ row_ptr,
                            const int32_t* __restrict__ gcol,
                            const double* __restrict__ w,
                            const double* __restrict__ cand_full,
                            const double* __restrict__ vol, double beta,
                            double beta_eta, double tbe, double denom,
                            const double* __restrict__ st,
                            double* __restrict__ part) {
  __shared__ double sh[kBlock / kWave];
  if (st[5] != 0.0) return;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int sub = threadIdx.x & (kSeg - 1);
  double vsum = 0.0;
  for (int64_t r = tid / kSeg; r < n_design; r += ((int64_t)gridDim.x * blockDim.x) / kSeg) {
    const int64_t s0 = row_ptr[r], s1 = row_ptr[r + 1];
    double acc = 0.0;
    for (int64_t q = s0 + sub; q < s1; q += kSeg) acc += w[q] * cand_full[gcol[q]];
#pragma unroll
    for (int off = kSeg / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kSeg);
    if (sub == 0) {
      double t = (acc - beta_eta) * beta;
      t = tanh(t);
      double proj = (tbe + t) / denom;
      vsum += proj * vol[r];
    }
  }
  double b = block_sum(vsum, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ void k_oc_volume_stencil(int n_design, const int32_t* __restrict__ ids, int nx,
                                    int ny, int nz, int R, int n_off,
                                    const int32_t* __restrict__ off,
                                    const int32_t* __restrict__ del,
                                    const double* __restrict__ u,
                                    const int32_t* __restrict__ colmap,
                                    const double* __restrict__ inv_rs,
                                    const double* __restrict__ cand_full,
                                    const double* __restrict__ vol, double beta,
                                    double beta_eta, double tbe, double denom,
                                    const double* __restrict__ st,
                                    double* __restrict__ part) {
  __shared__ double sh[kBlock / kWave];
  if (st[5] != 0.0) return;
  double vsum = 0.0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n_design;
       r += gridDim.x * blockDim.x) {
    const double acc = stencil_row<false, true>(ids[r], r, nx, ny, nz, R, n_off, off, del, u,
                                          colmap, inv_rs, cand_full);
    double t = (acc - beta_eta) * beta;
    t = tanh(t);
    double proj = (tbe + t) / denom;
    vsum += proj * vol[r];
  }
  double b = block_sum(vsum, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ void k_oc_bracket(int nb, const double* __restrict__ part,
                             double vol_sum, double vol_frac, double vol_tol,
                             double tolerance, int max_iter,
                             double* __restrict__ st) {
  __shared__ double sh[kBlock / kWave];
  if (st[5] != 0.0) return;
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) v += part[i];
  double tot = block_sum(v, sh);
  if (threadIdx.x == 0) {
    const double vol_error = tot / vol_sum - vol_frac;
    st[3] = vol_error;
    st[6] += 1.0;
    double l1 = st[0], l2 = st[1];
    const double it = st[4];
    if (fabs(vol_error) < vol_tol || it >= (double)max_iter ||
        fabs(l2 - l1) <= tolerance) {
      st[5] = 1.0;
    } else {
      if (vol_error > 0) l1 = st[2]; else l2 = st[2];
      st[0] = l1;
      st[1] = l2;
      st[4] = it + 1.0;
      st[2] = 0.5 * (l1 + l2);
    }
  }
}

// The bisection tree below bracket (l1, l2), heap order: node n's midpoint
// 0.5 (lo + hi), its children 2n + 1 (volume <= 0: l2 = mid) and 2n + 2
// (volume > 0: l1 = mid) -- the values the sequential updates compute.
__device__ __forceinline__ void oc_tree(double l1, double l2, double* mid) {
  double lo[kTreeK], hi[kTreeK];
  lo[0] = l1;
  hi[0] = l2;
#pragma unroll
  for (int n = 0; n < kTreeK; ++n) {
    mid[n] = 0.5 * (lo[n] + hi[n]);
    const int a = 2 * n + 1, b = 2 * n + 2;
    if (a < kTreeK) {
      lo[a] = lo[n];
      hi[a] = mid[n];
    }
    if (b < kTreeK) {
      lo[b] = mid[n];
      hi[b] = hi[n];
    }
  }
}

__global__ void k_oc_candidate_tree(int64_t n_design, const int32_t* __restrict__ ids,
                                    const double* __restrict__ rho_e,
                                    const double* __restrict__ dC,
                                    const double* __restrict__ st, double eps, double eta,
                                    double rate_min, double rate_max, double move_limit,
                                    double rho_min, double rho_max,
                                    double* __restrict__ cand_tree) {
  if (st[5] != 0.0) return;
  double mid[kTreeK];
  oc_tree(st[0], st[1], mid);
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_design;
       r += (int64_t)gridDim.x * blockDim.x) {
    double* out = cand_tree + (int64_t)ids[r] * kTreeStride;
#pragma unroll
    for (int k = 0; k < kTreeK; ++k)
      out[k] = oc_candidate(dC[r], rho_e[r], mid[k], eps, eta, rate_min, rate_max, move_limit,
                            rho_min, rho_max);
  }
}

// k_oc_volume at every midpoint of the tree: one read of the matrix, the
// same per-row and per-block summation order for each midpoint.
__global__ void k_oc_volume_tree(int64_t n_design, const int64_t* __restrict__ row_ptr,
                                 const int32_t* __restrict__ gcol,
                                 const double* __restrict__ w,
                                 const double* __restrict__ cand_tree,
                                 const double* __restrict__ vol, double beta, double beta_eta,
                                 double tbe, double denom, const double* __restrict__ st,
                                 double* __restrict__ part) {
  __shared__ double sh[kBlock / kWave];
  if (st[5] != 0.0) return;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int sub = threadIdx.x & (kSeg - 1);
  double vsum[kTreeK];
#pragma unroll
  for (int k = 0; k < kTreeK; ++k) vsum[k] = 0.0;
  for (int64_t r = tid / kSeg; r < n_design; r += ((int64_t)gridDim.x * blockDim.x) / kSeg) {
    const int64_t s0 = row_ptr[r], s1 = row_ptr[r + 1];
    double acc[kTreeK];
#pragma unroll
    for (int k = 0; k < kTreeK; ++k) acc[k] = 0.0;
    for (int64_t q = s0 + sub; q < s1; q += kSeg) {
      const double wq = w[q];
      const double* c = cand_tree + (int64_t)gcol[q] * kTreeStride;
#pragma unroll
      for (int k = 0; k < kTreeK; ++k) acc[k] += wq * c[k];
    }
#pragma unroll
    for (int k = 0; k < kTreeK; ++k)
#pragma unroll
      for (int off = kSeg / 2; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off, kSeg);
    if (sub == 0) {
#pragma unroll
      for (int k = 0; k < kTreeK; ++k) {
        double t = (acc[k] - beta_eta) * beta;
        t = tanh(t);
        double proj = (tbe + t) / denom;
        vsum[k] += proj * vol[r];
      }
    }
  }
  for (int k = 0; k < kTreeK; ++k) {
    const double b = block_sum(vsum[k], sh);
    if (threadIdx.x == 0) part[(int64_t)k * gridDim.x + blockIdx.x] = b;
    __syncthreads();
  }
}

// Walk the tree's path as the sequential bisection would (k_oc_bracket per
// level), stopping when it terminates or after max_evals evaluations in all;
// st[7] keeps the last evaluated midpoint.
__global__ void k_oc_bracket_tree(int nb, const double* __restrict__ part, double vol_sum,
                                  double vol_frac, double vol_tol, double tolerance,
                                  int max_iter, int max_evals, double* __restrict__ st) {
  __shared__ double sh[kBlock / kWave];
  __shared__ double tot[kTreeK];
  if (st[5] != 0.0) return;
  for (int k = 0; k < kTreeK; ++k) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) v += part[(int64_t)k * nb + i];
    const double t = block_sum(v, sh);
    if (threadIdx.x == 0) tot[k] = t;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double mid[kTreeK];
    oc_tree(st[0], st[1], mid);
    int node = 0;
    for (int lev = 0; lev < kTreeLevels; ++lev) {
      if (st[6] >= (double)max_evals) break;
      const double vol_error = tot[node] / vol_sum - vol_frac;
      st[3] = vol_error;
      st[6] += 1.0;
      st[7] = mid[node];
      double l1 = st[0], l2 = st[1];
      const double it = st[4];
      if (fabs(vol_error) < vol_tol || it >= (double)max_iter || fabs(l2 - l1) <= tolerance) {
        st[5] = 1.0;
        break;
      }
      if (vol_error > 0) {
        l1 = mid[node];
        node = 2 * node + 2;
      } else {
        l2 = mid[node];
        node = 2 * node + 1;
      }
      st[0] = l1;
      st[1] = l2;
      st[4] = it + 1.0;
      st[2] = 0.5 * (l1 + l2);
    }
  }
}

// the design rows' candidate at the last evaluated midpoint (st[7])
__global__ void k_oc_candidate_at(int64_t n_design, const double* __restrict__ rho_e,
                                  const double* __restrict__ dC,
                                  const double* __restrict__ st, double eps, double eta,
                                  double rate_min, double rate_max, double move_limit,
                                  double rho_min, double rho_max, double* __restrict__ rho_out) {
  const double lmid = st[7];
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_design;
       r += (int64_t)gridDim.x * blockDim.x)
    rho_out[r] = oc_candidate(dC[r], rho_e[r], lmid, eps, eta, rate_min, rate_max, move_limit,
                              rho_min, rho_max);
}

__global__ void k_oc_result(const double* __restrict__ st, double* __restrict__ res) {
  res[0] = st[2];
  res[1] = st[3];
  res[2] = st[6];
  res[3] = st[1] - st[0];
}

}  // namespace skt

using namespace skt;

template <typename F>
static int cub_call(F f, hipStream_t stream) {
  size_t bytes = 0;
  SKT_HIP(f(nullptr, bytes));
  void* tmp = nullptr;
  SKT_HIP(hipMalloc(&tmp, bytes ? bytes : 16));
  hipError_t e = f(tmp, bytes);
  hipError_t e2 = hipStreamSynchronize(stream);
  (void)hipFree(tmp);
  if (e != hipSuccess || e2 != hipSuccess) {
    set_error(std::string("hipcub: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    return SKT_EHIP;
  }
  return SKT_OK;
}

extern "C" {

int skt_filter_create(int64_t n_el, const double* centers,
                      const uint8_t* design, double sigma, double support,
                      skt_filter** out) {
  SKT_CHECK(out && centers && design, "null pointer");
  SKT_CHECK(n_el >= 1 && n_el < (int64_t)INT32_MAX, "n_el out of range");
  SKT_CHECK(sigma > 0.0 && support > 0.0, "sigma and support must be positive");
  *out = nullptr;
  hipStream_t stream = nullptr;
  skt_filter* f = new skt_filter();
  f->n_el = n_el;
  f->sigma = sigma;
  f->support = support;
  int st = SKT_OK;
  // --- bounding box on the host (setup only)
  std::vector<double> hc(3 * n_el);
  SKT_HIP(hipMemcpy(hc.data(), centers, sizeof(double) * 3 * n_el, hipMemcpyDeviceToHost));
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = hi[d] = hc[d * n_el];
    for (int64_t i = 0; i < n_el; ++i) {
      double v = hc[d * n_el + i];
      lo[d] = v < lo[d] ? v : lo[d];
      hi[d] = v > hi[d] ? v : hi[d];
    }
  }
  // cells at least the support wide, capped so the cell table stays small
  double cs = support * 1.000001;
  int64_t g[3];
  for (;;) {
    for (int d = 0; d < 3; ++d) g[d] = (int64_t)std::floor((hi[d] - lo[d]) / cs) + 1;
    if (g[0] * g[1] * g[2] <= 4 * n_el + 64) break;
    cs *= 1.5;
  }
  const int64_t ncell = g[0] * g[1] * g[2];
  int64_t *cell = nullptr, *cell_sorted = nullptr, *cstart = nullptr, *cend = nullptr;
  int32_t *ids = nullptr, *ids_sorted = nullptr, *flag = nullptr, *pos = nullptr;
  int64_t* counts = nullptr;
  SKT_HIP(hipMalloc(&cell, sizeof(int64_t) * n_el));
  SKT_HIP(hipMalloc(&cell_sorted, sizeof(int64_t) * n_el));
  SKT_HIP(hipMalloc(&ids, sizeof(int32_t) * n_el));
  SKT_HIP(hipMalloc(&ids_sorted, sizeof(int32_t) * n_el));
  SKT_HIP(hipMalloc(&cstart, sizeof(int64_t) * ncell));
  SKT_HIP(hipMalloc(&cend, sizeof(int64_t) * ncell));
  SKT_HIP(hipMemset(cstart, 0, sizeof(int64_t) * ncell));
  SKT_HIP(hipMemset(cend, 0, sizeof(int64_t) * ncell));
  SKT_HIP(hipMalloc(&flag, sizeof(int32_t) * n_el));
  SKT_HIP(hipMalloc(&pos, sizeof(int32_t) * n_el));
  const unsigned gb = grid_for(n_el);
  hipLaunchKernelGGL(k_cell_of, dim3(gb), dim3(kBlock), 0, stream, n_el, centers,
                     lo[0], lo[1], lo[2], 1.0 / cs, g[0], g[1], g[2], cell, ids);
  SKT_LAUNCH_CHECK();
  st = cub_call([&](void* t, size_t& b) {
    return hipcub::DeviceRadixSort::SortPairs(t, b, cell, cell_sorted, ids, ids_sorted,
                                              (int)n_el, 0, 64, stream);
  }, stream);
  if (st) return st;
  hipLaunchKernelGGL(k_cell_bounds, dim3(gb), dim3(kBlock), 0, stream, n_el,
                     cell_sorted, cstart, cend);
  hipLaunchKernelGGL(k_design_flags, dim3(gb), dim3(kBlock), 0, stream, n_el, design, flag);
  SKT_LAUNCH_CHECK();
  st = cub_call([&](void* t, size_t& b) {
    return hipcub::DeviceScan::ExclusiveSum(t, b, flag, pos, (int)n_el, stream);
  }, stream);
  if (st) return st;
  int32_t last_pos = 0, last_flag = 0;
  SKT_HIP(hipMemcpy(&last_pos, pos + n_el - 1, sizeof(int32_t), hipMemcpyDeviceToHost));
  SKT_HIP(hipMemcpy(&last_flag, flag + n_el - 1, sizeof(int32_t), hipMemcpyDeviceToHost));
  const int64_t nd = (int64_t)last_pos + last_flag;
  SKT_CHECK(nd >= 1, "design_elements is empty");
  f->n_design = nd;
  SKT_HIP(hipMalloc(&f->design_ids, sizeof(int32_t) * nd));
  SKT_HIP(hipMalloc(&f->colmap, sizeof(int32_t) * n_el));
  hipLaunchKernelGGL(k_design_ids, dim3(gb), dim3(kBlock), 0, stream, n_el, design,
                     pos, f->design_ids, f->colmap);
  SKT_LAUNCH_CHECK();
  SKT_HIP(hipMalloc(&counts, sizeof(int64_t) * (nd + 1)));
  SKT_HIP(hipMalloc(&f->row_ptr, sizeof(int64_t) * (nd + 1)));
  SKT_HIP(hipMemset(counts, 0, sizeof(int64_t) * (nd + 1)));
  const double r2 = support * support;
  const unsigned gd = grid_for(nd);
  hipLaunchKernelGGL(k_neighbours<false>, dim3(gd), dim3(kBlock), 0, stream, n_el, nd,
                     centers, f->design_ids, f->colmap, cell, ids_sorted, cstart, cend,
                     g[0], g[1], g[2], r2, counts, (const int64_t*)nullptr, (int32_t*)nullptr);
  SKT_LAUNCH_CHECK();
  st = cub_call([&](void* t, size_t& b) {
    return hipcub::DeviceScan::ExclusiveSum(t, b, counts, f->row_ptr, (int)(nd + 1), stream);
  }, stream);
  if (st) return st;
  int64_t nnz = 0;
  SKT_HIP(hipMemcpy(&nnz, f->row_ptr + nd, sizeof(int64_t), hipMemcpyDeviceToHost));
  SKT_CHECK(nnz >= 1, "No neighbor relations found; check support_radius or design_mask.");
  SKT_CHECK(nnz < (int64_t)1 << 40, "filter too large");
  f->nnz = nnz;
  SKT_HIP(hipMalloc(&f->gcol, sizeof(int32_t) * nnz));
  SKT_HIP(hipMalloc(&f->w, sizeof(double) * nnz));
  SKT_HIP(hipMalloc(&f->wt, sizeof(double) * nnz));
  SKT_HIP(hipMalloc(&f->inv_rs, sizeof(double) * nd));
  hipLaunchKernelGGL(k_neighbours<true>, dim3(gd), dim3(kBlock), 0, stream, n_el, nd,
                     centers, f->design_ids, f->colmap, cell, ids_sorted, cstart, cend,
                     g[0], g[1], g[2], r2, (int64_t*)nullptr, f->row_ptr, f->gcol);
  hipLaunchKernelGGL(k_sort_weights, dim3(gd), dim3(kBlock), 0, stream, n_el, nd,
                     centers, f->design_ids, f->row_ptr, f->gcol, sigma, f->w, f->inv_rs);
  hipLaunchKernelGGL(k_transpose_values, dim3(gd), dim3(kBlock), 0, stream, nd,
                     f->design_ids, f->colmap, f->row_ptr, f->gcol, f->w, f->wt);
  SKT_LAUNCH_CHECK();
  SKT_HIP(hipMalloc(&f->cand_full, sizeof(double) * n_el));
  SKT_HIP(hipMalloc(&f->cand_tree, sizeof(double) * kTreeStride * n_el));
  SKT_HIP(hipMalloc(&f->part_tree, sizeof(double) * kTreeK * kPartMax));
  SKT_HIP(hipMalloc(&f->part, sizeof(double) * kPartMax));
  SKT_HIP(hipMalloc(&f->state, sizeof(double) * 16));
  SKT_HIP(hipStreamSynchronize(stream));
  (void)hipFree(cell);
  (void)hipFree(cell_sorted);
  (void)hipFree(ids);
  (void)hipFree(ids_sorted);
  (void)hipFree(cstart);
  (void)hipFree(cend);
  (void)hipFree(flag);
  (void)hipFree(pos);
  (void)hipFree(counts);
  *out = f;
  return SKT_OK;
}

int skt_filter_set_grid(skt_filter* f, int64_t nx, int64_t ny, int64_t nz, double hx,
                        double hy, double hz, int* active) {
  SKT_CHECK(f && active, "null pointer");
  SKT_CHECK(nx >= 1 && ny >= 1 && nz >= 1 && nx * ny * nz == f->n_el,
            "grid does not match the filter's element count");
  SKT_CHECK(hx > 0.0 && hy > 0.0 && hz > 0.0, "grid spacing must be positive");
  *active = 0;
  f->stencil = 0;
  (void)hipFree(f->off);
  (void)hipFree(f->del);
  (void)hipFree(f->u_off);
  f->off = nullptr;
  f->del = nullptr;
  f->u_off = nullptr;
  SKT_CHECK(f->n_el < (int64_t)INT32_MAX, "grid too large for the stencil form");
  const double hmin = std::min(hx, std::min(hy, hz));
  const int R = (int)std::floor(f->support / hmin) + 1;
  if (R > 12) return SKT_OK;            // very wide stencils stay in CSR form
  const int wd = 2 * R + 1, nbox = wd * wd * wd;
  double* ubox = nullptr;
  int32_t* inside = nullptr;
  int32_t* bad = nullptr;
  SKT_HIP(hipMalloc(&ubox, sizeof(double) * nbox));
  SKT_HIP(hipMalloc(&inside, sizeof(int32_t) * nbox));
  SKT_HIP(hipMalloc(&bad, sizeof(int32_t)));
  SKT_HIP(hipMemset(bad, 0, sizeof(int32_t)));
  hipLaunchKernelGGL(k_offset_weights, dim3(grid_for(nbox)), dim3(kBlock), 0, nullptr, R, hx,
                     hy, hz, f->support * f->support, f->sigma, ubox, inside);
  hipLaunchKernelGGL(k_check_stencil, dim3(grid_for(f->n_design)), dim3(kBlock), 0, nullptr,
                     f->n_design, nx, ny, nz, R, f->design_ids, f->colmap, f->row_ptr,
                     f->gcol, f->w, f->inv_rs, ubox, inside, bad);
  SKT_LAUNCH_CHECK();
  std::vector<double> hu(nbox);
  std::vector<int32_t> hin(nbox);
  int32_t hbad = 1;
  SKT_HIP(hipMemcpy(hu.data(), ubox, sizeof(double) * nbox, hipMemcpyDeviceToHost));
  SKT_HIP(hipMemcpy(hin.data(), inside, sizeof(int32_t) * nbox, hipMemcpyDeviceToHost));
  SKT_HIP(hipMemcpy(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost));
  (void)hipFree(ubox);
  (void)hipFree(inside);
  (void)hipFree(bad);
  if (hbad) return SKT_OK;              // not the same matrix: keep the CSR form
  std::vector<int32_t> ho, hd;
  std::vector<double> hw;
  for (int dz = -R; dz <= R; ++dz)      // ascending element order of the neighbours
    for (int dy = -R; dy <= R; ++dy)
      for (int dx = -R; dx <= R; ++dx) {
        const int o = (dx + R) + wd * ((dy + R) + wd * (dz + R));
        if (!hin[o]) continue;
        ho.push_back((dx + 64) | ((dy + 64) << 8) | ((dz + 64) << 16));
        hd.push_back((int32_t)(dx + nx * (dy + ny * dz)));
        hw.push_back(hu[o]);
      }
  f->n_off = (int)hw.size();
  f->R = R;
  SKT_HIP(hipMalloc(&f->off, sizeof(int32_t) * ho.size()));
  SKT_HIP(hipMalloc(&f->del, sizeof(int32_t) * hd.size()));
  SKT_HIP(hipMalloc(&f->u_off, sizeof(double) * hw.size()));
  SKT_HIP(hipMemcpy(f->off, ho.data(), sizeof(int32_t) * ho.size(), hipMemcpyHostToDevice));
  SKT_HIP(hipMemcpy(f->del, hd.data(), sizeof(int32_t) * hd.size(), hipMemcpyHostToDevice));
  SKT_HIP(hipMemcpy(f->u_off, hw.data(), sizeof(double) * hw.size(), hipMemcpyHostToDevice));
  f->gnx = nx;
  f->gny = ny;
  f->gnz = nz;
  f->stencil = 1;
  // the stencil form's bisection evaluates one midpoint at a time
  (void)hipFree(f->cand_tree);
  (void)hipFree(f->part_tree);
  f->cand_tree = nullptr;
  f->part_tree = nullptr;
  *active = 1;
  return SKT_OK;
}

int skt_filter_destroy(skt_filter* f) {
  if (!f) return SKT_OK;
  (void)hipFree(f->design_ids);
  (void)hipFree(f->colmap);
  (void)hipFree(f->row_ptr);
  (void)hipFree(f->gcol);
  (void)hipFree(f->w);
  (void)hipFree(f->wt);
  (void)hipFree(f->inv_rs);
  (void)hipFree(f->off);
  (void)hipFree(f->del);
  (void)hipFree(f->u_off);
  (void)hipFree(f->cand_full);
  (void)hipFree(f->cand_tree);
  (void)hipFree(f->part_tree);
  (void)hipFree(f->part);
  (void)hipFree(f->state);
  delete f;
  return SKT_OK;
}

int skt_filter_info(const skt_filter* f, int64_t* n_design, int64_t* nnz) {
  SKT_CHECK(f, "null filter");
  if (n_design) *n_design = f->n_design;
  if (nnz) *nnz = f->nnz;
  return SKT_OK;
}

__global__ void k_export_cols(int64_t nnz, const int32_t* __restrict__ colmap,
                              const int32_t* __restrict__ gcol, int32_t* __restrict__ out) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz;
       q += (int64_t)gridDim.x * blockDim.x)
    out[q] = colmap[gcol[q]];
}

int skt_filter_export(const skt_filter* f, int64_t* indptr, int32_t* indices,
                      double* values, int32_t* design_ids, hipStream_t stream) {
  SKT_CHECK(f, "null filter");
  if (indptr)
    SKT_HIP(hipMemcpyAsync(indptr, f->row_ptr, sizeof(int64_t) * (f->n_design + 1),
                           hipMemcpyDeviceToDevice, stream));
  if (values)
    SKT_HIP(hipMemcpyAsync(values, f->w, sizeof(double) * f->nnz,
                           hipMemcpyDeviceToDevice, stream));
  if (design_ids)
    SKT_HIP(hipMemcpyAsync(design_ids, f->design_ids, sizeof(int32_t) * f->n_design,
                           hipMemcpyDeviceToDevice, stream));
  if (indices) {
    hipLaunchKernelGGL(k_export_cols, dim3(grid_for(f->nnz)), dim3(kBlock), 0, stream,
                       f->nnz, f->colmap, f->gcol, indices);
    SKT_LAUNCH_CHECK();
  }
  return SKT_OK;
}

int skt_filter_forward(skt_filter* f, const double* x, double* y, hipStream_t stream) {
  SKT_CHECK(f && x && y, "null pointer");
  SKT_CHECK(x != y, "filter forward cannot run in place");
  if (f->stencil)
    hipLaunchKernelGGL(k_stencil_filter<false>, dim3(grid_for(f->n_el)), dim3(kBlock), 0,
                       stream, (int)f->n_el, (int)f->gnx, (int)f->gny, (int)f->gnz, f->R,
                       f->n_off, f->off, f->del, f->u_off, f->colmap, f->inv_rs, x, y);
  else
    hipLaunchKernelGGL(k_csr_apply, dim3(grid_for(f->n_el * kSeg)), dim3(kBlock), 0, stream,
                       f->n_el, f->colmap, f->row_ptr, f->gcol, f->w, x, y, 0);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_filter_gradient(skt_filter* f, const double* v, double* g, hipStream_t stream) {
  SKT_CHECK(f && v && g, "null pointer");
  SKT_CHECK(v != g, "filter gradient cannot run in place");
  if (f->stencil)
    hipLaunchKernelGGL(k_stencil_filter<true>, dim3(grid_for(f->n_el)), dim3(kBlock), 0,
                       stream, (int)f->n_el, (int)f->gnx, (int)f->gny, (int)f->gnz, f->R,
                       f->n_off, f->off, f->del, f->u_off, f->colmap, f->inv_rs, v, g);
  else
    hipLaunchKernelGGL(k_csr_apply, dim3(grid_for(f->n_el * kSeg)), dim3(kBlock), 0, stream,
                       f->n_el, f->colmap, f->row_ptr, f->gcol, f->wt, v, g, 1);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_filter_gather_design(skt_filter* f, const double* full, double* design,
                             hipStream_t stream) {
  SKT_CHECK(f && full && design, "null pointer");
  hipLaunchKernelGGL(k_gather_design, dim3(grid_for(f->n_design)), dim3(kBlock), 0, stream,
                     f->n_design, f->design_ids, full, design);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

int skt_filter_scatter_design(skt_filter* f, const double* design, double* full,
                              hipStream_t stream) {
  SKT_CHECK(f && full && design, "null pointer");
  hipLaunchKernelGGL(k_scatter_design, dim3(grid_for(f->n_design)), dim3(kBlock), 0, stream,
                     f->n_design, f->design_ids, design, full);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

// SKT_OC_TREE=0 evaluates the CSR form's bisection one midpoint at a time
// (same decisions and results; read at every call)
static bool oc_tree_enabled() {
  const char* v = std::getenv("SKT_OC_TREE");
  return !(v && v[0] == '0' && v[1] == 0);
}

int skt_oc_bisect(skt_filter* f, const double* rho_full, const double* rho_e,
                  const double* dC, const double* vol, double vol_sum,
                  double vol_frac, double move_limit, double eta,
                  double rho_min, double rho_max, double rate_min,
                  double rate_max, double beta, double beta_eta, double l1,
                  double l2, double tolerance, double vol_tol, int max_iter,
                  double* rho_out, double* result, hipStream_t stream) {
  SKT_CHECK(f && rho_full && rho_e && dC && vol && rho_out && result, "null pointer");
  SKT_CHECK(l1 < l2, "lambda_lower must be < lambda_upper");
  SKT_CHECK(tolerance > 0.0 && max_iter >= 0, "bad bisection limits");
  // number of evaluations is bounded by the bracket halving to `tolerance`
  const double w0 = l2 - l1;
  int steps = (int)std::ceil(std::log2(w0 / tolerance)) + 3;
  if (steps < 1) steps = 1;
  if (steps > max_iter + 1) steps = max_iter + 1;
  const double tbe = std::tanh(beta * beta_eta);
  const double denom = std::tanh(beta * beta_eta) + std::tanh(beta * (1.0 - beta_eta)) + 1e-12;
  const int64_t nd = f->n_design;
  const unsigned nbs = grid_for(nd, kBlock, kPartMax);
  unsigned nb = grid_for(nd * kSeg, kBlock, kPartMax);
  // candidates live on the design rows only; the other entries stay zero,
  // which the stencil volume sum relies on (its design-masked row form)
  hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(1), 0, stream, f->state, l1, l2);
  if (!f->stencil && f->cand_tree && oc_tree_enabled()) {
    // CSR form: kTreeLevels bisection steps per read of the matrix; the
    // decisions, evaluation count and result equal the step-by-step loop's.
    // (The stencil form stays one midpoint per pass: its volume kernel reads
    // no matrix, and seven candidates per offset cost more than the launches
    // they save.)
    hipLaunchKernelGGL(k_fill, dim3(grid_for(kTreeStride * f->n_el)), dim3(kBlock), 0, stream,
                       kTreeStride * f->n_el, 0.0, f->cand_tree);
    for (int s = 0; s < steps; s += kTreeLevels) {
      hipLaunchKernelGGL(k_oc_candidate_tree, dim3(grid_for(nd)), dim3(kBlock), 0, stream, nd,
                         f->design_ids, rho_e, dC, f->state, 1e-12, eta, rate_min, rate_max,
                         move_limit, rho_min, rho_max, f->cand_tree);
      hipLaunchKernelGGL(k_oc_volume_tree, dim3(nb), dim3(kBlock), 0, stream, nd, f->row_ptr,
                         f->gcol, f->w, f->cand_tree, vol, beta, beta_eta, tbe, denom, f->state,
                         f->part_tree);
      hipLaunchKernelGGL(k_oc_bracket_tree, dim3(1), dim3(kBlock), 0, stream, (int)nb,
                         f->part_tree, vol_sum, vol_frac, vol_tol, tolerance, max_iter, steps,
                         f->state);
    }
    hipLaunchKernelGGL(k_oc_candidate_at, dim3(grid_for(nd)), dim3(kBlock), 0, stream, nd, rho_e,
                       dC, f->state, 1e-12, eta, rate_min, rate_max, move_limit, rho_min,
                       rho_max, rho_out);
    hipLaunchKernelGGL(k_oc_result, dim3(1), dim3(1), 0, stream, f->state, result);
    SKT_LAUNCH_CHECK();
    return SKT_OK;
  }
  hipLaunchKernelGGL(k_fill, dim3(grid_for(f->n_el)), dim3(kBlock), 0, stream, f->n_el, 0.0,
                     f->cand_full);
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(k_oc_candidate, dim3(grid_for(nd)), dim3(kBlock), 0, stream, nd,
                       f->design_ids, rho_e, dC, f->state, 1e-12, eta, rate_min, rate_max,
                       move_limit, rho_min, rho_max, f->cand_full, rho_out);
    if (f->stencil)
      hipLaunchKernelGGL(k_oc_volume_stencil, dim3(nbs), dim3(kBlock), 0, stream, (int)nd,
                         f->design_ids, (int)f->gnx, (int)f->gny, (int)f->gnz, f->R, f->n_off,
                         f->off, f->del, f->u_off, f->colmap, f->inv_rs, f->cand_full, vol,
                         beta, beta_eta, tbe, denom, f->state, f->part);
    else
      hipLaunchKernelGGL(k_oc_volume, dim3(nb), dim3(kBlock), 0, stream, nd, f->design_ids,
                         f->row_ptr, f->gcol, f->w, f->cand_full, vol, beta, beta_eta, tbe,
                         denom, f->state, f->part);
    hipLaunchKernelGGL(k_oc_bracket, dim3(1), dim3(kBlock), 0, stream,
                       (int)(f->stencil ? nbs : nb), f->part,
                       vol_sum, vol_frac, vol_tol, tolerance, max_iter, f->state);
  }
  hipLaunchKernelGGL(k_oc_result, dim3(1), dim3(1), 0, stream, f->state, result);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

}  // extern "C"
