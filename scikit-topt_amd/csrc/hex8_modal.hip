// Structured hex8 operator apply in the modal (Walsh-Hadamard) basis of the
// trilinear element, LDS-tiled and streamed along z.
//
// For an axis-aligned box element the element matrix becomes very sparse in
// the basis of the 8 corner sign patterns (1, x, y, z, xy, yz, xz, xyz):
// per element and column y_e = E_e * H (Kmod (H^T u_e)), with H the +-1
// butterfly (24 add/sub per component each way) and Kmod a 21x21 matrix with
// 33 distinct coefficients (elasticity) or an 8-entry diagonal (scalar:
// conduction, mass, or a combination),
// instead of a dense 24x24 (8x8) product.
//
// Tiling: a workgroup owns a 15x15 patch of nodes in x-y (16 along an axis
// for the last tile when exactly one node is left over: the missing 16th
// element of that row lies outside the grid) and a chunk of node planes in
// z.  It walks the element layers bottom-up: each step loads one
// new node plane (17x17 with halo) into LDS, 256 threads compute the 16x16
// elements of the layer that touch the patch, park the 8 corner results of
// each element in LDS, and every owned node sums its 4 contributions from
// the layer below and 4 from the layer above in a fixed order (no atomics,
// bit-reproducible).  Each node thread writes its completed values itself
// (contiguous per node in the column-blocked layout); Dirichlet rows
// (identity) are applied on the way out.  Two barriers per element layer.
//
// Vector layout: value (dof, column c) lives at
//   (c / lkc) * gstride + (c % lkc) + dof * dstride,
// which covers the public (ndof, k) row-major block (lkc = k, dstride = k)
// and the solver's column-blocked [k/kc][ndof][kc] layout (lkc = kc,
// gstride = ndof*kc, dstride = kc).  A workgroup handles KC columns; with
// the blocked layout its loads and stores are contiguous.
#include <cstdlib>

#include <hip/hip_ext.h>

#include "operator.hpp"

namespace skt {

constexpr int kTE = 16;          // elements per tile edge (one ghost element)
constexpr int kTN = kTE - 1;     // owned nodes per tile edge (16 on a closing tile)
constexpr int kPW = kTE + 1;     // node plane width incl. halo
constexpr int kPN = kPW * kPW;

// LDS slot of halo-plane node (lx, ly), lx, ly in [0, 17): columns 0..15 at
// lx + 16 ly, column 16 packed after them.  A half-wavefront reads two
// element rows (16 consecutive lx each, rows ly and ly + 1), which land in
// disjoint bank ranges with a 16-word row pitch.
__device__ __forceinline__ int node_slot(int lx, int ly) {
  return lx < 16 ? lx + 16 * ly : 16 * kPW + ly;
}

// Value of the next lane (i + 1) within each row of 16 lanes (DPP
// row_shl:1); lane 15 of a row receives 0.  Element lanes are laid out 16 per
// x-row, so this fetches the right-hand neighbour element's result.
__device__ __forceinline__ float from_right(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x101, 0xf, 0xf, false));
}
__device__ __forceinline__ double from_right(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x101, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x101, 0xf, 0xf, false);
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}

// 3-stage Walsh-Hadamard butterfly on 8 values indexed by corner bits
// (bit 0: x, bit 1: y, bit 2: z): out[S] = sum_i (-1)^popcount(i & S) v[i].
// The matrix is symmetric and squares to 8 I, so the same butterfly maps
// modal forces back to corners.
// 2-stage butterfly on the 4 corners of a node-plane face (bit 0: x, bit 1:
// y): the x-y part of wht8.
template <typename T>
__device__ __forceinline__ void wht4(T v[4]) {
  T p = v[0] + v[1], q = v[0] - v[1];
  T r = v[2] + v[3], t = v[2] - v[3];
  v[0] = p + r;
  v[2] = p - r;
  v[1] = q + t;
  v[3] = q - t;
}

template <typename T>
__device__ __forceinline__ void wht8(T v[8]) {
#pragma unroll
  for (int b = 1; b < 8; b <<= 1) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (i & b) continue;
      const T p = v[i] + v[i | b];
      const T q = v[i] - v[i | b];
      v[i] = p;
      v[i | b] = q;
    }
  }
}

// f = Kmod m for elasticity, modes indexed by bits (1=x 2=y 4=z 3=xy 6=yz
// 5=xz 7=xyz), m[c][S] component c.  Coefficient layout (see host packer):
// lin[6] (x0,y1,z2 sym: 00 11 22 01 02 12), sxy[3] {(x,1),(y,0)}: aa bb ab,
// sxz[3] {(x,2),(z,0)}, syz[3] {(y,2),(z,1)}, b1[3] {(xy,0),(yz,2)},
// b2[3] {(xy,1),(xz,2)}, b3[3] {(yz,1),(xz,0)},
// b4[6] {(xy,2),(yz,0),(xz,1)} sym: 00 11 22 01 02 12, t[3] (xyz,c).
template <typename T>
__device__ __forceinline__ void modal_elastic(const T* __restrict__ c, const T m[3][8],
                                              T f[3][8]) {
  {
    const T a = m[0][1], b = m[1][2], d = m[2][4];
    f[0][1] = c[0] * a + c[3] * b + c[4] * d;
    f[1][2] = c[3] * a + c[1] * b + c[5] * d;
    f[2][4] = c[4] * a + c[5] * b + c[2] * d;
  }
#define SKT_PAIR(o, ra, sa, rb, sb)                    \
  {                                                    \
    const T a = m[ra][sa], b = m[rb][sb];              \
    f[ra][sa] = c[o] * a + c[o + 2] * b;               \
    f[rb][sb] = c[o + 2] * a + c[o + 1] * b;           \
  }
  SKT_PAIR(6, 1, 1, 0, 2)    // sxy
  SKT_PAIR(9, 2, 1, 0, 4)    // sxz
  SKT_PAIR(12, 2, 2, 1, 4)   // syz
  SKT_PAIR(15, 0, 3, 2, 6)   // b1
  SKT_PAIR(18, 1, 3, 2, 5)   // b2
  SKT_PAIR(21, 1, 6, 0, 5)   // b3
#undef SKT_PAIR
  {
    const T a = m[2][3], b = m[0][6], d = m[1][5];
    f[2][3] = c[24] * a + c[27] * b + c[28] * d;
    f[0][6] = c[27] * a + c[25] * b + c[29] * d;
    f[1][5] = c[28] * a + c[29] * b + c[26] * d;
  }
  f[0][7] = c[30] * m[0][7];
  f[1][7] = c[31] * m[1][7];
  f[2][7] = c[32] * m[2][7];
  f[0][0] = f[1][0] = f[2][0] = T(0);
}

template <typename T>
__device__ __forceinline__ void modal_heat(const T* __restrict__ c, const T m[1][8], T f[1][8]) {
  // scalar operators are diagonal in the modal basis; mode 0 is 0 for pure
  // conduction and carries the mass of a*M + b*K
#pragma unroll
  for (int s = 0; s < 8; ++s) f[0][s] = c[s] * m[0][s];
}

template <typename T>
struct ModalArgs {
  T c[40];
};

// LDS row stride for a structure-of-arrays buffer of n entries holding
// `combos` interleaved (component, column) rows: odd and ~32/combos mod 32
// so the combos written by neighbouring lanes land in different banks.
constexpr int soa_stride(int n, int combos) {
  const int want = ((32 / combos) | 1) % 32;
  int s = n;
  while (s % 32 != want) ++s;
  return s;
}

// What the kernel writes for an owned value y = (K x)_i (see OutMode).
template <typename T>
struct OutArgs {
  const T* B = nullptr;        // right-hand side (same layout as X)
  const T* dinv = nullptr;     // (ndof) Jacobi inverse
  const double* cheb = nullptr;  // (c1, c2) of this step, device
  int prev_zero = 0;           // the iterate before X is zero (first smoothing step)
  const double* cheb0 = nullptr;  // OUT_CHEB0: coefficients of the implied first step
  // DOT kernels: per-workgroup partial sums of the fused dot product,
  // dot[c * dot_nb + bx] for column c and workgroup bx of its column block
  double* dot = nullptr;
  int dot_nb = 0;
  // column stop flags (kStopStride apart): a workgroup whose columns have all
  // stopped returns at once
  const double* stop = nullptr;
};

// SF (sum-factorised), bit 0: the x-y butterfly of each element's face on a
// node plane is computed once per plane and carried in registers to the
// next layer (the element thread keeps its footprint as the z-stream
// advances), so a layer reads 4 corners instead of 8 and does 16 instead of
// 24 add/sub per component forward.  Bit 1: the backward butterfly
// accumulates the face forces of the layer below and this layer in modal
// space and finishes each node plane with one 2-D butterfly, leaving half
// the pair traffic.  The carries cost registers, so which form is fastest
// depends on the precision and output mode.
//
// DOT: the kernel also leaves the per-column partial sums of x . y
// (OUT_APPLY: the CG's p . K p) or b . x_out (OUT_CHEB: r . z of a V-cycle's
// last smoothing step) over its owned nodes, reduced in a fixed order.
template <typename T, int NC, int KC, int MODE, int SF, bool DOT>
__global__ void __launch_bounds__(256)
k_hex_apply_modal(int64_t nx, int64_t ny, int64_t nz, int lkc, int64_t gstride, int dstride,
                  int lz, int tiles_x, int tiles_y, int n_zchunks, ModalArgs<T> coef,
                  const T* __restrict__ E, const uint8_t* __restrict__ fixed,
                  const T* __restrict__ X, T* __restrict__ Y, OutArgs<T> oa) {
  // LDS, structure-of-arrays so that consecutive lanes touch consecutive
  // words: two rotating node planes [NC*KC][kPS]; per element lane the
  // x-paired corner sums [side][cy][NC][KC][256] (side 0: node plane below the
  // element, 1: above; cy: node row).
  constexpr int kPS = soa_stride(kPN, NC * KC);
  constexpr int kPlane = kPS * NC * KC;
  constexpr int kItems = kPN * NC * KC;
  constexpr int kPre = (kItems + 255) / 256;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* planes = reinterpret_cast<T*>(smem_raw);
  T* pairs = planes + 2 * kPlane;
  const int64_t nnx = nx + 1, nny = ny + 1, nnz = nz + 1;
  const int64_t plane_stride = nnx * nny * NC * (int64_t)dstride;
  int b = blockIdx.x;
  const int tx = b % tiles_x; b /= tiles_x;
  const int ty = b % tiles_y; b /= tiles_y;
  const int zc = b % n_zchunks; b /= n_zchunks;
  const int c0 = b * KC;
  if (oa.stop) {
    bool all = true;
#pragma unroll
    for (int c = 0; c < KC; ++c) all = all && oa.stop[(c0 + c) * kStopStride] != 0.0;
    if (all) return;
  }
  const int64_t cbase = (int64_t)(c0 / lkc) * gstride + (c0 % lkc);
  X += cbase;
  Y += cbase;
  constexpr bool kCheb = MODE == OUT_CHEB || MODE == OUT_CHEB0;
  constexpr bool kC0 = MODE == OUT_CHEB0;
  // float64 OUT_CHEB0 reads b and dinv of the owned nodes in the epilogue
  // rather than at the start of the layer (158 instead of 178 VGPRs, three
  // waves per SIMD instead of two); in float32 the early loads fit in 128
  constexpr bool kLate = kC0 && sizeof(T) == 8;
  const T* Bv = MODE != OUT_APPLY ? oa.B + cbase : nullptr;
  T c1 = T(0), c2 = T(0), s0 = T(0);
  if (kCheb) {
    c1 = (T)oa.cheb[0];
    c2 = (T)oa.cheb[1];
  }
  if (MODE == OUT_CHEB0) s0 = (T)oa.cheb0[1];
  // the previous iterate is read (own values only) from the output buffer
  // it is about to be replaced in
  const bool use_prev = MODE == OUT_CHEB && c1 != T(0) && !oa.prev_zero;
  const int64_t x0 = (int64_t)tx * kTN, y0 = (int64_t)ty * kTN;
  const int64_t z0 = (int64_t)zc * lz;
  const int64_t z1 = (z0 + lz < nnz) ? z0 + lz : nnz;
  const int tid = threadIdx.x;
  const int ex_l = tid % kTE, ey_l = tid / kTE;
  const int64_t ex = x0 - 1 + ex_l, ey = y0 - 1 + ey_l;
  const bool el_in = ex >= 0 && ex < nx && ey >= 0 && ey < ny;
  const int64_t e_base = el_in ? ex + nx * ey : 0;
  // owned nodes: kTN per axis, or all remaining (at most kTE) on the last tile
  const int own_x = (int)(tx == tiles_x - 1 ? nnx - x0 : kTN);
  const int own_y = (int)(ty == tiles_y - 1 ? nny - y0 : kTN);
  const int nx_l = tid % kTE, ny_l = tid / kTE;
  const bool is_node = nx_l < own_x && ny_l < own_y;
  const int e_lo = tid;  // element lane holding row ny_l's pair (x0 - 1 + nx_l, y0 - 1 + ny_l)
  const bool has_up = ny_l + 1 < kTE;  // element row ny_l + 1 exists in the tile
  const int own_slot = node_slot(nx_l + 1, ny_l + 1);  // owned node in the halo plane
  const int64_t own_node = (y0 + ny_l) * nnx + (x0 + nx_l);  // in-plane index
  // the owned node's values: ((own_node * NC + d) * dstride + col) in a plane
  const int64_t own_off = own_node * NC * (int64_t)dstride;

  // staged item r of this thread: LDS slot, and its in-plane dof (node * NC +
  // comp; -1 outside the grid); the column of every item is tid % KC (256 is a
  // multiple of KC), so the value offset is dof * dstride + tid % KC.
  // OUT_CHEB0 keeps only the dof and forms the offset at each fetch.
  int f_off[kC0 ? 1 : kPre], f_lds[kPre];
  int f_dof[kC0 ? kPre : 1];
  const int f_col = tid % KC;
#pragma unroll
  for (int r = 0; r < kPre; ++r) {
    const int i = tid + 256 * r;
    f_off[kC0 ? 0 : r] = -1;
    f_lds[r] = -1;
    if constexpr (kC0) f_dof[r] = -1;
    if (i < kItems) {
      const int col = i % KC, comp = (i / KC) % NC, nl = i / (KC * NC);
      const int64_t gx = x0 - 1 + nl % kPW, gy = y0 - 1 + nl / kPW;
      f_lds[r] = (comp * KC + col) * kPS + node_slot(nl % kPW, nl / kPW);
      if (gx >= 0 && gx < nnx && gy >= 0 && gy < nny) {
        if constexpr (kC0)
          f_dof[r] = (int)((gy * nnx + gx) * NC + comp);
        else
          f_off[r] = (int)(((gy * nnx + gx) * NC + comp) * dstride + col);
      }
    }
  }

  // prefetched plane values stay in flight until stash (the compute of the
  // current layer runs meanwhile); OUT_CHEB0 fetches dinv separately, after
  // the layer's compute (fetch_dinv), and forms the implied first iterate
  // s0 dinv b at stash time
  T pre[kPre];
  T pred[kC0 ? kPre : 1];
  // The loads are unpredicated and unselected: an item outside the grid
  // reads offset 0 of the plane, and stash writes zero for it.  A predicated
  // load, or a select right after the load, compiles to a branch with a wait
  // for that load inside it, which drains every load in flight.
  auto fetch = [&](int64_t gz) {
    const T* Xp = X + (gz >= 0 && gz < nnz ? gz : 0) * plane_stride;
#pragma unroll
    for (int r = 0; r < kPre; ++r) {
      if constexpr (kC0)
        pre[r] = Xp[(f_dof[r] >= 0 ? f_dof[r] : 0) * dstride + f_col];
      else
        pre[r] = Xp[f_off[r] >= 0 ? f_off[r] : 0];
    }
  };
  auto fetch_dinv = [&](int64_t gz) {
    if constexpr (kC0) {
      const T* Dp = oa.dinv + (gz >= 0 && gz < nnz ? gz : 0) * nnx * nny * NC;
#pragma unroll
      for (int r = 0; r < kPre; ++r) pred[r] = Dp[f_dof[r] >= 0 ? f_dof[r] : 0];
    }
  };
  auto stash = [&](T* dst, int64_t gz) {
    const bool zin = gz >= 0 && gz < nnz;
#pragma unroll
    for (int r = 0; r < kPre; ++r) {
      if (f_lds[r] < 0) continue;
      const bool ok = zin && (kC0 ? f_dof[r] : f_off[kC0 ? 0 : r]) >= 0;
      if constexpr (MODE == OUT_CHEB0)
        dst[f_lds[r]] = ok ? s0 * pred[r] * pre[r] : T(0);
      else
        dst[f_lds[r]] = ok ? pre[r] : T(0);
    }
  };

  T partial[KC][NC];
#pragma unroll
  for (int c = 0; c < KC; ++c)
#pragma unroll
    for (int d = 0; d < NC; ++d) partial[c][d] = T(0);
  double dacc[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) dacc[c] = 0.0;

  // Dirichlet flags of the owned node on plane ez (zero-offset reads where
  // the thread writes nothing on that plane)
  uint8_t fxn[NC];
  auto load_flags = [&](int64_t ez) {
    const int64_t pn = (is_node && ez >= z0 && ez < z1) ? (ez * nnx * nny + own_node) * NC : 0;
#pragma unroll
    for (int d = 0; d < NC; ++d) fxn[d] = fixed[pn + d];
  };
  load_flags(z0 - 1);

  int bot = 0, top = 1;
  fetch(z0 - 1);
  fetch_dinv(z0 - 1);
  stash(planes + bot * kPlane, z0 - 1);
  fetch(z0);
  fetch_dinv(z0);
  stash(planes + top * kPlane, z0);
  // SF carries: face butterfly of the element's footprint on the bottom
  // plane, and the top-face forces of the layer below
  constexpr bool kSFF = (SF & 1) != 0, kSFB = (SF & 2) != 0;
  T Fb[kSFF ? KC : 1][NC][4], Gc[kSFB ? KC : 1][NC][4];
  if constexpr (kSFF) {
    __syncthreads();
#pragma unroll
    for (int col = 0; col < KC; ++col)
#pragma unroll
      for (int d = 0; d < NC; ++d) {
#pragma unroll
        for (int i2 = 0; i2 < 4; ++i2)
          Fb[col][d][i2] = planes[bot * kPlane + (d * KC + col) * kPS +
                                  node_slot(ex_l + (i2 & 1), ey_l + (i2 >> 1))];
        wht4(Fb[col][d]);
      }
  }
  if constexpr (kSFB) {
#pragma unroll
    for (int col = 0; col < KC; ++col)
#pragma unroll
      for (int d = 0; d < NC; ++d)
#pragma unroll
        for (int i2 = 0; i2 < 4; ++i2) Gc[col][d][i2] = T(0);
  }
  for (int64_t ez = z0 - 1; ez < z1; ++ez) {
    // this layer's operands are loaded before the prefetch of plane ez + 2, so
    // that waiting for them leaves the prefetch in flight; all unpredicated
    // (clamped addresses; a value a thread does not use is never selected)
    const T Ev = E[e_base + nx * ny * (ez < 0 ? 0 : (ez < nz ? ez : nz - 1))];
    // this step completes node plane ez; its owned values are written by the
    // node threads, whose extra operands are loaded now so that their latency
    // hides behind the layer's compute
    const bool out = is_node && ez >= z0;
    const int64_t po = ez * plane_stride + own_off;
    const int64_t po_c = out ? po : 0;
    const int64_t pn_c = out ? (ez * nnx * nny + own_node) * NC : 0;
    // the Dirichlet flags come one layer ahead: loaded in the epilogue, where
    // they are used, they would be issued after the prefetch and their wait
    // would drain it
    uint8_t fxv[NC];
#pragma unroll
    for (int d = 0; d < NC; ++d) fxv[d] = fxn[d];
    load_flags(ez + 1);
    // (kLate: loaded in the epilogue instead)
    T ob[KC][NC], op[KC][NC], oi[NC];
#pragma unroll
    for (int d = 0; d < NC; ++d) {
      if constexpr (kCheb && !kLate) oi[d] = oa.dinv[pn_c + d];
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        if constexpr (MODE != OUT_APPLY && !kLate) ob[c][d] = Bv[po_c + d * dstride + c];
        // without a previous iterate every thread reads offset 0
        if constexpr (MODE == OUT_CHEB && !kLate)
          op[c][d] = Y[use_prev ? po_c + d * dstride + c : 0];
      }
    }
    fetch(ez + 2);  // in flight while this layer is computed
    __syncthreads();
    const T Ee = (el_in && ez >= 0 && ez < nz) ? Ev : T(0);
    const T* pb = planes + bot * kPlane;
    const T* pt = planes + top * kPlane;
#pragma unroll
    for (int col = 0; col < KC; ++col) {
      T m[NC][8];
      if constexpr (kSFF) {
        // z stage from the carried bottom face and this top face
#pragma unroll
        for (int d = 0; d < NC; ++d) {
          T ft[4];
#pragma unroll
          for (int i2 = 0; i2 < 4; ++i2)
            ft[i2] = pt[(d * KC + col) * kPS + node_slot(ex_l + (i2 & 1), ey_l + (i2 >> 1))];
          wht4(ft);
#pragma unroll
          for (int i2 = 0; i2 < 4; ++i2) {
            m[d][i2] = (Fb[col][d][i2] + ft[i2]) * Ee;
            m[d][i2 | 4] = (Fb[col][d][i2] - ft[i2]) * Ee;
            Fb[col][d][i2] = ft[i2];
          }
        }
      } else {
#pragma unroll
        for (int idx = 0; idx < 8; ++idx) {
          const int cx = idx & 1, cy = (idx >> 1) & 1, cz = idx >> 2;
          const int nl = node_slot(ex_l + cx, ey_l + cy);
          const T* p = cz ? pt : pb;
#pragma unroll
          for (int d = 0; d < NC; ++d) m[d][idx] = p[(d * KC + col) * kPS + nl] * Ee;
        }
#pragma unroll
        for (int d = 0; d < NC; ++d) wht8(m[d]);
      }
      T f[NC][8];
      if constexpr (NC == 3) modal_elastic(coef.c, m, f);
      else modal_heat(coef.c, m, f);
      if constexpr (kSFB) {
        // bottom-face forces of this layer plus top-face forces of the layer
        // below complete node plane ez; the top-face forces are carried
#pragma unroll
        for (int d = 0; d < NC; ++d) {
          T g[4];
#pragma unroll
          for (int i2 = 0; i2 < 4; ++i2) {
            g[i2] = Gc[col][d][i2] + (f[d][i2] + f[d][i2 | 4]);
            Gc[col][d][i2] = f[d][i2] - f[d][i2 | 4];
          }
          wht4(g);
          // node (ex+1, ey+cy) gets this element's corner cx=1 plus the
          // right-hand element's corner cx=0
#pragma unroll
          for (int cy = 0; cy < 2; ++cy) {
            const T v = g[1 | (cy << 1)] + from_right(g[cy << 1]);
            pairs[((cy * NC + d) * KC + col) * 256 + tid] = v;
          }
        }
      } else {
#pragma unroll
        for (int d = 0; d < NC; ++d) wht8(f[d]);
        // node (ex+1, ey+cy) of plane side cz gets this element's corner cx=1
        // plus the right-hand element's corner cx=0
#pragma unroll
        for (int q = 0; q < 4; ++q) {  // q = cy | cz << 1
          const int cy = q & 1, cz = q >> 1;
#pragma unroll
          for (int d = 0; d < NC; ++d) {
            const T v = f[d][1 | (cy << 1) | (cz << 2)] + from_right(f[d][(cy << 1) | (cz << 2)]);
            pairs[(((cz * 2 + cy) * NC + d) * KC + col) * 256 + tid] = v;
          }
        }
      }
    }
    // the owned node's x (node plane ez, the bottom LDS plane), for Dirichlet
    // rows (y = x) and the Chebyshev update: read after the compute, so that
    // it holds no registers there, and before the barrier, after which other
    // waves' stash overwrites that plane
    T xo[KC][NC];
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
      for (int d = 0; d < NC; ++d) xo[c][d] = pb[(d * KC + c) * kPS + own_slot];
    __syncthreads();
    fetch_dinv(ez + 2);  // for the stash below
    if (is_node) {
      T* Yp = Y + po;
#pragma unroll
      for (int col = 0; col < KC; ++col)
#pragma unroll
        for (int d = 0; d < NC; ++d) {
          // element row ny_l + 1 is outside the tile only for the 16th node
          // row of a closing tile, where it is outside the grid as well
          const bool fixd = out && fxv[d] != 0;
          T y;
          if constexpr (kSFB) {
            const T tot = pairs[((1 * NC + d) * KC + col) * 256 + e_lo] +
                          (has_up ? pairs[((0 * NC + d) * KC + col) * 256 + e_lo + kTE] : T(0));
            y = fixd ? xo[col][d] : tot;
          } else {
            const T below = pairs[(((0 * 2 + 1) * NC + d) * KC + col) * 256 + e_lo] +
                            (has_up ? pairs[(((0 * 2 + 0) * NC + d) * KC + col) * 256 + e_lo + kTE]
                                    : T(0));
            const T above = pairs[(((1 * 2 + 1) * NC + d) * KC + col) * 256 + e_lo] +
                            (has_up ? pairs[(((1 * 2 + 0) * NC + d) * KC + col) * 256 + e_lo + kTE]
                                    : T(0));
            y = fixd ? xo[col][d] : partial[col][d] + below;
            partial[col][d] = above;
          }
          if (out) {
            const int64_t o = d * dstride + col;
            if constexpr (MODE == OUT_APPLY) {
              Yp[o] = y;
              if constexpr (DOT) dacc[col] += (double)xo[col][d] * (double)y;
            } else if constexpr (MODE == OUT_RESID) {
              // zero at Dirichlet DOFs, which the restriction drops
              Yp[o] = fixd ? T(0) : ob[col][d] - y;
            } else {
              if constexpr (kLate) ob[col][d] = Bv[po + o];
              // Chebyshev step with the direction recovered from the two
              // iterates: x_out = x + c1 (x - x_prev) + c2 dinv (b - K x)
              // (x_prev = 0 for the implied first step of OUT_CHEB0)
              T xp = (MODE == OUT_CHEB && use_prev) ? op[col][d] : T(0);
              T di = oi[d];
              if constexpr (kLate) {
                if constexpr (MODE == OUT_CHEB) xp = use_prev ? Yp[o] : T(0);
                di = oa.dinv[(ez * nnx * nny + own_node) * NC + d];
              }
              const T dn = c1 * (xo[col][d] - xp) + c2 * di * (ob[col][d] - y);
              const T xn = xo[col][d] + dn;
              Yp[o] = xn;
              if constexpr (DOT) dacc[col] += (double)ob[col][d] * (double)xn;
            }
          }
        }
    }
    // the bottom plane receives plane ez + 2 and becomes the new top (every
    // read of it in this step happened before the second barrier)
    stash(planes + bot * kPlane, ez + 2);
    const int t = bot;
    bot = top;
    top = t;
  }
  if constexpr (DOT) {
    // wavefront sums, then the four wavefronts in order; the pair area is
    // free once every node thread has read it
    double* red = reinterpret_cast<double*>(pairs);
    const int lane = tid & (kWave - 1), w = tid / kWave;
#pragma unroll
    for (int c = 0; c < KC; ++c) dacc[c] = wave_sum(dacc[c]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < KC; ++c) red[w * KC + c] = dacc[c];
    }
    __syncthreads();
    if (tid < KC) {
      double v = 0.0;
      for (int i = 0; i < 256 / kWave; ++i) v += red[i * KC + tid];
      oa.dot[(int64_t)(c0 + tid) * oa.dot_nb + (int)(blockIdx.x % (unsigned)oa.dot_nb)] = v;
    }
  }
}

// Element energies 0.5 E_e u_e^T K0 u_e in the modal basis: with m = H u
// per component (H the 8-point butterfly above, symmetric) and f = Kmod m,
// the element apply is H f, so u^T K0 u = m . f.  One thread per (element,
// column), columns fastest so a node's columns are read together.
template <typename T, int NC>
__global__ void k_hex_energy_modal(int64_t nx, int64_t ny, int64_t nz, int ncols,
                                   ModalArgs<T> coef, const T* __restrict__ E,
                                   const T* __restrict__ U, double* __restrict__ energy) {
  const int64_t nnx = nx + 1, nny = ny + 1;
  const int64_t total = nx * ny * nz * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int col = (int)(t % ncols);
    const int64_t e = t / ncols;
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    T m[NC][8];
#pragma unroll
    for (int idx = 0; idx < 8; ++idx) {
      const int64_t nb = (ex + (idx & 1)) + nnx * ((ey + ((idx >> 1) & 1)) + nny * (ez + (idx >> 2)));
#pragma unroll
      for (int d = 0; d < NC; ++d) m[d][idx] = U[(nb * NC + d) * ncols + col];
    }
#pragma unroll
    for (int d = 0; d < NC; ++d) wht8(m[d]);
    T f[NC][8];
    if constexpr (NC == 3) modal_elastic(coef.c, m, f);
    else modal_heat(coef.c, m, f);
    double q = 0.0;
#pragma unroll
    for (int d = 0; d < NC; ++d)
#pragma unroll
      for (int i = 0; i < 8; ++i) q += (double)m[d][i] * (double)f[d][i];
    energy[e * ncols + col] = 0.5 * (double)E[e] * q;
  }
}

template <typename T>
int hex_energy_modal(const skt_operator* op, int ncols, const T* coef, const T* E, const T* U,
                     double* out, hipStream_t s) {
  ModalArgs<T> a;
  for (int i = 0; i < 40; ++i) a.c[i] = coef[i];
  const int64_t total = op->n_el * ncols;
  if (op->ncomp == 3)
    hipLaunchKernelGGL((k_hex_energy_modal<T, 3>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                       op->nx, op->ny, op->nz, ncols, a, E, U, out);
  else
    hipLaunchKernelGGL((k_hex_energy_modal<T, 1>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                       op->nx, op->ny, op->nz, ncols, a, E, U, out);
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

template int hex_energy_modal<float>(const skt_operator*, int, const float*, const float*,
                                     const float*, double*, hipStream_t);
template int hex_energy_modal<double>(const skt_operator*, int, const double*, const double*,
                                      const double*, double*, hipStream_t);

// Launch geometry: (tiles_x, tiles_y) node tiles per plane, nzc z chunks of
// lz node planes, one workgroup per (tile, chunk, column block) with the
// column block slowest-varying.
struct ModalGrid {
  int tiles_x, tiles_y, nzc;
  int64_t lz, nblocks;
};

static ModalGrid modal_grid(const skt_operator* op, int k, int kc_block) {
  ModalGrid g;
  const int64_t nnx = op->nx + 1, nny = op->ny + 1, nnz = op->nz + 1;
  // the last tile of an axis takes up to kTE nodes, so ceil((nn - 1) / kTN)
  g.tiles_x = (int)((nnx - 1 + kTN - 1) / kTN);
  g.tiles_y = (int)((nny - 1 + kTN - 1) / kTN);
  const int cblocks = k / kc_block;
  // z chunk length: aim for >= 4 workgroups per CU (256 CUs), but keep the
  // one-layer warm-up of each chunk below ~1/6 of its work
  g.lz = nnz;
  const int64_t per = (int64_t)g.tiles_x * g.tiles_y * cblocks;
  while (g.lz > 6 && per * ((nnz + g.lz - 1) / g.lz) < 1024) g.lz = (g.lz + 1) / 2;
  g.nzc = (int)((nnz + g.lz - 1) / g.lz);
  g.nblocks = per * g.nzc;
  return g;
}

// Columns per workgroup for a layout with column group kc: two for
// elasticity (more columns cut LDS occupancy below three workgroups per CU),
// up to four for heat.
static int modal_kc(const skt_operator* op, int kc) {
  if (op->ncomp == 3) return kc % 2 == 0 ? 2 : 1;
  return kc % 4 == 0 ? 4 : (kc % 2 == 0 ? 2 : 1);
}

int hex_modal_dot_slots(const skt_operator* op, int k, int kc) {
  const ModalGrid g = modal_grid(op, k, modal_kc(op, kc));
  return g.tiles_x * g.tiles_y * g.nzc;
}

// Events for the next modal launch on this host thread (modal_time_next):
// the launch records them itself through hipExtLaunchKernelGGL.
static thread_local hipEvent_t g_time_start = nullptr, g_time_stop = nullptr;

void modal_time_next(hipEvent_t start, hipEvent_t stop) {
  g_time_start = start;
  g_time_stop = stop;
}

bool modal_time_clear() {
  const bool pending = g_time_start != nullptr;
  g_time_start = g_time_stop = nullptr;
  return pending;
}

// Launch helper for structured grids whose element matrix has the modal
// pattern.  k columns in the layout described above.
template <typename T, int NC, int KC, int MODE, int SF, bool DOT>
static int launch_modal(const skt_operator* op, int k, int lkc, int64_t gstride, int dstride,
                        const T* coef, const T* E, const T* X, T* Y, const OutArgs<T>& oa,
                        hipStream_t s) {
  SKT_CHECK((op->nx + 1) * (op->ny + 1) * NC * (int64_t)dstride < INT32_MAX,
            "node plane too large for 32-bit in-plane offsets");
  SKT_CHECK(k % KC == 0 && lkc % KC == 0, "column blocking does not divide the layout");
  ModalArgs<T> a;
  for (int i = 0; i < 40; ++i) a.c[i] = coef[i];
  const ModalGrid gr = modal_grid(op, k, KC);
  if (DOT) {
    // column blocks are the slowest-varying workgroup index, so workgroup
    // b % (nblocks / cblocks) is the partial's slot within its column
    SKT_CHECK(oa.dot && oa.dot_nb == gr.tiles_x * gr.tiles_y * gr.nzc,
              "internal: dot partials do not match the launch");
  }
  const size_t shmem = sizeof(T) * ((size_t)2 * soa_stride(kPN, NC * KC) * NC * KC +
                                    (size_t)((SF & 2) ? 2 : 4) * NC * KC * 256);
  static bool attr_set = false;
  if (!attr_set) {
    SKT_HIP(hipFuncSetAttribute((const void*)k_hex_apply_modal<T, NC, KC, MODE, SF, DOT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    attr_set = true;
  }
  // the flags (all zero without fixed DOFs) are read unconditionally
  SKT_CHECK(op->fixed != nullptr, "internal: hex8 operator without Dirichlet flags");
  const uint8_t* fx = op->fixed;
  if (g_time_start) {
    // timed launch: the events take the dispatch's own start and end
    hipEvent_t t0 = g_time_start, t1 = g_time_stop;
    g_time_start = g_time_stop = nullptr;
    hipExtLaunchKernelGGL((k_hex_apply_modal<T, NC, KC, MODE, SF, DOT>),
                          dim3((unsigned)gr.nblocks), dim3(256), (uint32_t)shmem, s, t0, t1, 0u,
                          op->nx, op->ny, op->nz, lkc, gstride, dstride, (int)gr.lz, gr.tiles_x,
                          gr.tiles_y, gr.nzc, a, E, fx, X, Y, oa);
  } else {
    hipLaunchKernelGGL((k_hex_apply_modal<T, NC, KC, MODE, SF, DOT>),
                       dim3((unsigned)gr.nblocks), dim3(256), shmem, s, op->nx, op->ny, op->nz,
                       lkc, gstride, dstride, (int)gr.lz, gr.tiles_x, gr.tiles_y, gr.nzc, a, E,
                       fx, X, Y, oa);
  }
  SKT_LAUNCH_CHECK();
  return SKT_OK;
}

// Butterfly form per precision and output mode: the carried (sum-factorised)
// form pays off where the extra registers still leave two waves per SIMD --
// the float64 plain apply (118 against 129 us at 120x60x60, 8 columns);
// float32 and the fused residual / Chebyshev outputs keep the element-wise
// butterflies.  SKT_MODAL_SF=0/1 forces one form everywhere.
template <typename T, int MODE>
static bool modal_sf() {
  static const int forced = [] {
    const char* v = std::getenv("SKT_MODAL_SF");
    return (v && (v[0] == '0' || v[0] == '1') && v[1] == 0) ? v[0] - '0' : -1;
  }();
  if (forced >= 0) return forced == 1;
  return sizeof(T) == 8 && MODE == OUT_APPLY;
}

// Butterfly form and fused dot product of one (NC, KC) instance.
template <typename T, int MODE, int NC, int KC>
static int launch_sel(const skt_operator* op, int k, int lkc, int64_t gstride, int dstride,
                      const T* coef, const T* E, const T* X, T* Y, const OutArgs<T>& oa,
                      hipStream_t s) {
  const bool sf = modal_sf<T, MODE>();
  if (oa.dot) {
    if constexpr (MODE == OUT_APPLY || MODE == OUT_CHEB) {
      return sf ? launch_modal<T, NC, KC, MODE, 3, true>(op, k, lkc, gstride, dstride, coef, E,
                                                         X, Y, oa, s)
                : launch_modal<T, NC, KC, MODE, 0, true>(op, k, lkc, gstride, dstride, coef, E,
                                                         X, Y, oa, s);
    } else {
      set_error("internal: no fused dot product for this output mode");
      return SKT_EINVAL;
    }
  }
  return sf ? launch_modal<T, NC, KC, MODE, 3, false>(op, k, lkc, gstride, dstride, coef, E, X,
                                                      Y, oa, s)
            : launch_modal<T, NC, KC, MODE, 0, false>(op, k, lkc, gstride, dstride, coef, E, X,
                                                      Y, oa, s);
}

template <typename T, int MODE>
static int dispatch_modal(const skt_operator* op, int k, int kc, const T* coef, const T* E,
                          const T* X, T* Y, const OutArgs<T>& oa, hipStream_t s) {
  // kc == k: public (ndof, k) layout; otherwise column-blocked with group kc
  const int lkc = kc;
  const int64_t gstride = (kc == k) ? 0 : op->ndof * kc;
  const int dstride = (kc == k) ? k : kc;
#define SKT_LM(NC_, KC_) \
  return launch_sel<T, MODE, NC_, KC_>(op, k, lkc, gstride, dstride, coef, E, X, Y, oa, s)
  const int kb = modal_kc(op, kc);
  if (op->ncomp == 3) {
    if (kb == 2) SKT_LM(3, 2);
    SKT_LM(3, 1);
  }
  if (kb == 4) SKT_LM(1, 4);
  if (kb == 2) SKT_LM(1, 2);
  SKT_LM(1, 1);
#undef SKT_LM
}

template <typename T>
int hex_apply_modal(const skt_operator* op, int k, int kc, const T* coef, const T* E, const T* X,
                    T* Y, hipStream_t s, double* dot, const double* stop) {
  OutArgs<T> oa;
  oa.stop = stop;
  if (dot) {
    oa.dot = dot;
    oa.dot_nb = hex_modal_dot_slots(op, k, kc);
  }
  return dispatch_modal<T, OUT_APPLY>(op, k, kc, coef, E, X, Y, oa, s);
}

template <typename T>
int hex_resid_modal(const skt_operator* op, int k, int kc, const T* coef, const T* E,
                    const T* X, const T* B, T* R, hipStream_t s, const double* stop) {
  OutArgs<T> oa;
  oa.stop = stop;
  oa.B = B;
  return dispatch_modal<T, OUT_RESID>(op, k, kc, coef, E, X, R, oa, s);
}

template <typename T>
int hex_cheb0_modal(const skt_operator* op, int k, int kc, const T* coef, const T* E, const T* B,
                    T* Xout, const T* dinv, const double* cheb0, const double* cheb,
                    hipStream_t s, const double* stop) {
  OutArgs<T> oa;
  oa.stop = stop;
  oa.B = B;
  oa.dinv = dinv;
  oa.cheb = cheb;
  oa.cheb0 = cheb0;
  oa.prev_zero = 1;
  return dispatch_modal<T, OUT_CHEB0>(op, k, kc, coef, E, B, Xout, oa, s);
}

template <typename T>
int hex_cheb_modal(const skt_operator* op, int k, int kc, const T* coef, const T* E, const T* X,
                   T* Xout, const T* B, const T* dinv, const double* cheb, int prev_zero,
                   hipStream_t s, double* dot, const double* stop) {
  OutArgs<T> oa;
  oa.stop = stop;
  oa.B = B;
  oa.dinv = dinv;
  oa.cheb = cheb;
  oa.prev_zero = prev_zero;
  if (dot) {
    oa.dot = dot;
    oa.dot_nb = hex_modal_dot_slots(op, k, kc);
  }
  return dispatch_modal<T, OUT_CHEB>(op, k, kc, coef, E, X, Xout, oa, s);
}

template int hex_apply_modal<float>(const skt_operator*, int, int, const float*, const float*,
                                    const float*, float*, hipStream_t, double*, const double*);
template int hex_apply_modal<double>(const skt_operator*, int, int, const double*, const double*,
                                     const double*, double*, hipStream_t, double*,
                                     const double*);
template int hex_resid_modal<float>(const skt_operator*, int, int, const float*, const float*,
                                    const float*, const float*, float*, hipStream_t,
                                    const double*);
template int hex_resid_modal<double>(const skt_operator*, int, int, const double*,
                                     const double*, const double*, const double*, double*,
                                     hipStream_t, const double*);
template int hex_cheb0_modal<float>(const skt_operator*, int, int, const float*, const float*,
                                    const float*, float*, const float*, const double*,
                                    const double*, hipStream_t, const double*);
template int hex_cheb0_modal<double>(const skt_operator*, int, int, const double*,
                                     const double*, const double*, double*, const double*,
                                     const double*, const double*, hipStream_t, const double*);
template int hex_cheb_modal<float>(const skt_operator*, int, int, const float*, const float*,
                                   const float*, float*, const float*, const float*,
                                   const double*, int, hipStream_t, double*, const double*);
template int hex_cheb_modal<double>(const skt_operator*, int, int, const double*, const double*,
                                    const double*, double*, const double*, const double*,
                                    const double*, int, hipStream_t, double*, const double*);

}  // namespace skt
