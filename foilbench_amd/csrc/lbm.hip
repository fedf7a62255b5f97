// D2Q9 two-relaxation-time LBM on gfx950.
//
// State layout (SoA, x fastest): the persistent field is the POST-collision population
// f*[d][y][x] -- exactly the array the reference materialises between its collision and
// streaming kernels (kern/lbm.py:12-76 -> 80-148). One fused kernel per lattice step
//   pull-stream f* (periodic index wrap, kern/lbm.py:270-274)
//   -> interpolated moving-wall bounce-back at cut links (kern/lbm.py:276-308)
//   -> Zou-He / channel rows, inlet, convective outlet, corners (fb/solvers/lbm.py:612-645)
//   -> sponge (kern/lbm.py:152-168)
//   -> [uncovered-cell reset for the next pose (fb/solvers/lbm.py:463-484)]
//   -> TRT collision (kern/lbm.py:12-76) -> write f*,
// so each lattice step reads 9 and writes 9 populations per cell. The pre-collision field
// (the reference's solver state) is written only on the last substep of an advance, where
// the postcondition reductions are fused in.
#include "lbm_batch.h"
#include <type_traits>
#include <vector>

#include "diag.h"
#include "geometry.h"

namespace fb {

enum : uint8_t { FL_SOLID = 1, FL_NEAR = 2 };

struct D2Q9 {
  static constexpr int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr double w[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

template <int P>
struct Prec;
// FP64 storage, FP64 arithmetic.
template <>
struct Prec<FB_F64> {
  using S = double;
  using C = double;
  using B = double;
  static constexpr bool shift = false;
};
// FP32 storage, FP64 kernel arithmetic, FP32 boundary arithmetic (the reference's float32).
template <>
struct Prec<FB_F32> {
  using S = float;
  using C = double;
  using B = float;
  static constexpr bool shift = false;
};
// Shifted FP32 storage (f - w), FP64 collision arithmetic: with FP32 arithmetic the collision's rounding is
// at the scale of h but amplified by the relaxation, and dominates the field error (1024 x 512 lattice, 20000
// steps: 1.3e-5 against FP64, with FP64 arithmetic 9e-7, for 2% more time per step).
template <>
struct Prec<FB_F32_MIXED> {
  using S = float;
  using C = double;
  using B = float;
  static constexpr bool shift = true;
};
// Shifted FP16 storage (f - w), FP32 arithmetic.
template <>
struct Prec<FB_F16_MIXED> {
  using S = __half;
  using C = float;
  using B = float;
  static constexpr bool shift = true;
};

template <class T>
__device__ __forceinline__ T ldv(const double* p) { return (T)*p; }

__device__ __forceinline__ double to_c(double v) { return v; }
__device__ __forceinline__ float to_cf(float v) { return v; }

template <class S, class C>
__device__ __forceinline__ C ld(const S* p) { return (C)(*p); }
template <>
__device__ __forceinline__ float ld<__half, float>(const __half* p) { return __half2float(*p); }

template <class S, class C>
__device__ __forceinline__ void st(S* p, C v) { *p = (S)v; }
template <>
__device__ __forceinline__ void st<__half, float>(__half* p, float v) { *p = __float2half_rn(v); }

// Streaming forms of the population loads and stores (NT): nontemporal, so a lattice larger than the
// caches streams through them without evicting what the step reuses (flags, near-wall data).
template <bool NT, class S>
__device__ __forceinline__ S ld_pop(const S* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <>
__device__ __forceinline__ __half ld_pop<true, __half>(const __half* p) {
  return __ushort_as_half(__builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p)));
}
template <bool NT, class S, class C>
__device__ __forceinline__ void st_pop(S* p, C v) {
  if constexpr (NT) __builtin_nontemporal_store((S)v, p);
  else *p = (S)v;
}
template <>
__device__ __forceinline__ void st_pop<true, __half, float>(__half* p, float v) {
  __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(v)), reinterpret_cast<unsigned short*>(p));
}

// Round a compute value to the storage grid (only for the reference-faithful forms).
template <class P>
__device__ __forceinline__ typename P::C rnd(typename P::C v) {
  if constexpr (P::shift) return v;
  else return (typename P::C)(typename P::S)v;
}

struct LbmFeq {
  double v[9];
};

template <class S>
struct LbmArgs {
  int nx, ny;
  const S* __restrict__ fin;      // post-collision f* of the previous step (or pre-collision f for priming)
  S* __restrict__ fout;           // post-collision f* of this step
  S* __restrict__ fpre;           // pre-collision f of this step (written when write_pre)
  const S* __restrict__ rho_in;   // collision density at near-wall cells of the streaming geometry
  S* __restrict__ rho_out;        // collision density at near-wall cells of the next geometry
  const S* __restrict__ outlet_in;
  S* __restrict__ outlet_out;
  const uint8_t* __restrict__ flags;       // streaming geometry
  const uint8_t* __restrict__ flags_next;  // geometry of the next collision
  const double* __restrict__ sdf;
  const double* __restrict__ sponge_y;
  const double* __restrict__ sponge_x;
  double x0, y0, dx, dy, pivot_x, pivot_y;
  double wall_omega, wall_scale;
  double even, odd;
  double u_lattice, tux, tuy;
  double vscale;
  double feq_b[9];
  double feq_u[9];
  int periodic_x, periodic_y, channel;
  int write_pre, do_reset, prime;
  // launch shape: blocks [0, edge_blocks) cover the edge cells (edge_index), the next list_blocks
  // stride over the flagged (solid or near-wall) cells of rows 1..ny-2 listed in `list`;
  // the rest cover rows 1..ny-2, int_rows rows per block (1 or 2) and int_gx blocks across,
  // whose edge and flagged lanes idle
  int int_gx, int_rows, edge_blocks, list_blocks;
  int skip;  // batched launches: this lattice sits out the step (its blocks return at once)
  const int* __restrict__ list;
  const int* __restrict__ list_count;
  const double* __restrict__ links;  // 8 Bouzidi fractions per list entry (0: straight link)
  const int* __restrict__ lidx;      // list position of each listed cell (other cells: unset)
  unsigned long long* red;  // [min f key, max speed key, max excursion key, nonfinite f, bad rho, nonfinite u]
  // per-block slots of the statistics ([6][blocks]), folded by lbm_stats_fold_kernel; null: each wave merges
  // into `red` with atomics (batched launches)
  unsigned long long* spart;
};

// Where a lattice step reads the previous post-collision state f*, the collision density of the
// near-wall cells and the outlet history, and where it writes the outlet history it produces.
// PlaneSrc: the global planes. TileSrc: a block's LDS tile holding the first step of a fused
// pair over the block's cells and a one-cell halo.
template <class S>
struct PlaneSrc {
  const S* __restrict__ f;
  const S* __restrict__ rho;
  const S* __restrict__ hist;
  S* hist_out;
  int hist_y0;  // hist_out holds rows hist_y0.. (0 for the global history)
  int nx;
  size_t N;
  __device__ __forceinline__ S at(int d, int x, int y) const { return f[(size_t)d * N + (size_t)y * nx + x]; }
  __device__ __forceinline__ S at_in(int d, int x, int y) const { return at(d, x, y); }
  __device__ __forceinline__ S rho_at(int x, int y) const { return rho[(size_t)y * nx + x]; }
  __device__ __forceinline__ S hist_at(int y, int d) const { return hist[(size_t)y * 9 + d]; }
  __device__ __forceinline__ void put_hist(int y, int d, S v) const { hist_out[(size_t)(y - hist_y0) * 9 + d] = v; }
};

// LDS tile of W x H cells whose corner is the domain cell (ox, oy). A read outside the tile is a
// periodic-wrap read of a non-periodic lattice, whose value the boundary treatment overwrites
// (Zou-He unknowns, corner equilibria); it returns 0.
template <class S, int W, int H>
struct TileSrc {
  const S* f;     // [9][H][W]
  const S* rho;   // [H][W]
  const S* hist;  // [H][9]: outlet history written by the first step, tile rows
  S* hist_out;    // global outlet history of the second step
  int ox, oy;
  __device__ __forceinline__ S at(int d, int x, int y) const {
    const unsigned lx = (unsigned)(x - ox), ly = (unsigned)(y - oy);
    return (lx < (unsigned)W && ly < (unsigned)H) ? f[d * (W * H) + (int)ly * W + (int)lx] : (S)0.0f;
  }
  // a read known to fall inside the tile
  __device__ __forceinline__ S at_in(int d, int x, int y) const { return f[d * (W * H) + (y - oy) * W + (x - ox)]; }
  __device__ __forceinline__ S rho_at(int x, int y) const { return rho[(y - oy) * W + (x - ox)]; }
  __device__ __forceinline__ S hist_at(int y, int d) const { return hist[(y - oy) * 9 + d]; }
  __device__ __forceinline__ void put_hist(int y, int d, S v) const { hist_out[(size_t)y * 9 + d] = v; }
};

// Interior blocks of the single-step kernel with `rows` interior rows per block.
__host__ __device__ inline int interior_blocks(int nx, int ny, int rows) {
  return (nx + 255) / 256 * ((ny - 2 + rows - 1) / rows);
}

template <class S>
__device__ __forceinline__ PlaneSrc<S> plane_src(const LbmArgs<S>& a) {
  return PlaneSrc<S>{a.fin, a.rho_in, a.outlet_in, a.outlet_out, 0, a.nx, (size_t)a.nx * a.ny};
}

// ------------------------------------------------------------------ streaming pieces

#pragma clang fp contract(off)

// Bouzidi reflection at fluid cell (x,y) for the link e pointing into solid cell n;
// the result lands in direction opp(e) at (x,y) (kern/lbm.py:276-308).
template <class P, class Src>
__device__ __forceinline__ typename P::C reflect(const LbmArgs<typename P::S>& a, const Src& src_f, int x, int y,
                                                 int e, int nxs, int nys) {
  using S = typename P::S;
  using C = typename P::C;
  const size_t i = (size_t)y * a.nx + x;
  double sx = a.sdf[i];
  double sn = a.sdf[(size_t)nys * a.nx + nxs];
  double den = fmax(sx - sn, 1.0e-12);
  double q = fmin(fmax(sx / den, 0.05), 1.0);
  const S sv = src_f.at(e, x, y);
  C src = ld<S, C>(&sv);
  C r;
  if (q < 0.5) {
    int ux = wrapi(x - D2Q9::cx[e], a.nx), uy = wrapi(y - D2Q9::cy[e], a.ny);
    const S uv = src_f.at(e, ux, uy);
    C up = ld<S, C>(&uv);
    C q2 = (C)(2.0 * q);
    C q12 = (C)(1.0 - 2.0 * q);
    r = q2 * src + q12 * up;
  } else {
    C q2 = (C)(2.0 * q);
    C q21 = (C)(2.0 * q - 1.0);
    const S ov = src_f.at(D2Q9::opp[e], x, y);
    C fo = ld<S, C>(&ov);
    r = src / q2 + q21 * fo / q2;
  }
  if (a.wall_omega != 0.0) {
    double wx = a.x0 + ((double)x + 0.5) * a.dx + (q * (double)D2Q9::cx[e]) * a.dx;
    double wy = a.y0 + ((double)y + 0.5) * a.dy + (q * (double)D2Q9::cy[e]) * a.dy;
    double wu = (-a.wall_omega) * (wy - a.pivot_y);
    double wv = a.wall_omega * (wx - a.pivot_x);
    double wp = (double)D2Q9::cx[e] * wu + (double)D2Q9::cy[e] * wv;
    const S rv = src_f.rho_at(x, y);
    C rho = ld<S, C>(&rv);
    if constexpr (P::shift) rho = rho + (C)1;
    r = r - (C)6.0 * (C)D2Q9::w[e] * rho * (C)wp * (C)a.wall_scale;
  }
  return rnd<P>(r);
}

// Pull-stream of cell (x,y): the value each direction receives (kern/lbm.py:262-308).
template <class P, class Src>
__device__ __forceinline__ void pull(const LbmArgs<typename P::S>& a, const Src& src, int x, int y,
                                     typename P::C g[9]) {
  using S = typename P::S;
  using C = typename P::C;
  const size_t i = (size_t)y * a.nx + x;
  const S v0 = src.at(0, x, y);
  g[0] = ld<S, C>(&v0);
  if (a.flags[i] & FL_SOLID) {
#pragma unroll
    for (int d = 1; d < 9; ++d) g[d] = (C)0;
    if constexpr (P::shift) {
#pragma unroll
      for (int d = 1; d < 9; ++d) g[d] = -(C)D2Q9::w[d];
    }
    return;
  }
  const bool near = a.flags[i] & FL_NEAR;
#pragma unroll
  for (int d = 1; d < 9; ++d) {
    int sx = wrapi(x - D2Q9::cx[d], a.nx), sy = wrapi(y - D2Q9::cy[d], a.ny);
    size_t s = (size_t)sy * a.nx + sx;
    if (near && (a.flags[s] & FL_SOLID)) {
      g[d] = reflect<P>(a, src, x, y, D2Q9::opp[d], sx, sy);
    } else {
      const S v = src.at(d, sx, sy);
      g[d] = ld<S, C>(&v);
    }
  }
}

// Pull-stream with every straight source loaded up front: cells whose own flags are clear
// (no solid cell among the neighbours) need nothing else; flagged cells fall back to pull().
template <class P, class Src>
__device__ __forceinline__ void pull_fast(const LbmArgs<typename P::S>& a, const Src& src, int x, int y,
                                          typename P::C g[9]) {
  using S = typename P::S;
  using C = typename P::C;
  const size_t i = (size_t)y * a.nx + x;
  S v[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    const int sx = wrapi(x - D2Q9::cx[d], a.nx), sy = wrapi(y - D2Q9::cy[d], a.ny);
    v[d] = src.at(d, sx, sy);
  }
  if (a.flags[i] != 0) {
    pull<P>(a, src, x, y, g);
    return;
  }
#pragma unroll
  for (int d = 0; d < 9; ++d) g[d] = ld<S, C>(&v[d]);
}

// Pull-stream of a flagged interior cell from its precomputed links: q[d-1] > 0 marks the
// direction d whose source cell is solid, holding the Bouzidi fraction reflect() would compute
// (same expression, so the same bits). Every operand is loaded up front, in one round.
template <class P, class Src>
__device__ __forceinline__ void pull_linked(const LbmArgs<typename P::S>& a, const Src& src, int x, int y,
                                            bool solid, const double* __restrict__ q, typename P::C g[9]) {
  using S = typename P::S;
  using C = typename P::C;
  if (solid) {
    const S v0 = src.at(0, x, y);
    g[0] = ld<S, C>(&v0);
#pragma unroll
    for (int d = 1; d < 9; ++d) g[d] = P::shift ? -(C)D2Q9::w[d] : (C)0;
    return;
  }
  // operands stay in the storage type until used, to keep the register footprint small
  S at[9], straight[9], up[9];
  double qd[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) at[d] = src.at(d, x, y);
#pragma unroll
  for (int d = 1; d < 9; ++d) {
    const int sx = wrapi(x - D2Q9::cx[d], a.nx), sy = wrapi(y - D2Q9::cy[d], a.ny);
    straight[d] = src.at(d, sx, sy);
    // upstream of the reflected link e = opp(d): the cell (x - c_e) = (x + c_d)
    const int ux = wrapi(x + D2Q9::cx[d], a.nx), uy = wrapi(y + D2Q9::cy[d], a.ny);
    up[d] = src.at(D2Q9::opp[d], ux, uy);
    qd[d] = q[d - 1];
  }
  C rho_w = (C)0;
  if (a.wall_omega != 0.0) {
    const S rv = src.rho_at(x, y);
    rho_w = ld<S, C>(&rv);
    if constexpr (P::shift) rho_w = rho_w + (C)1;
  }
  g[0] = ld<S, C>(&at[0]);
#pragma unroll
  for (int d = 1; d < 9; ++d) {
    const double qq = qd[d];
    if (!(qq > 0.0)) {
      g[d] = ld<S, C>(&straight[d]);
      continue;
    }
    const int e = D2Q9::opp[d];
    const C src = ld<S, C>(&at[e]);
    C r;
    if (qq < 0.5) {
      C q2 = (C)(2.0 * qq);
      C q12 = (C)(1.0 - 2.0 * qq);
      r = q2 * src + q12 * ld<S, C>(&up[d]);
    } else {
      C q2 = (C)(2.0 * qq);
      C q21 = (C)(2.0 * qq - 1.0);
      r = src / q2 + q21 * ld<S, C>(&at[d]) / q2;
    }
    if (a.wall_omega != 0.0) {
      double wx = a.x0 + ((double)x + 0.5) * a.dx + (qq * (double)D2Q9::cx[e]) * a.dx;
      double wy = a.y0 + ((double)y + 0.5) * a.dy + (qq * (double)D2Q9::cy[e]) * a.dy;
      double wu = (-a.wall_omega) * (wy - a.pivot_y);
      double wv = a.wall_omega * (wx - a.pivot_x);
      double wp = (double)D2Q9::cx[e] * wu + (double)D2Q9::cy[e] * wv;
      r = r - (C)6.0 * (C)D2Q9::w[e] * rho_w * (C)wp * (C)a.wall_scale;
    }
    g[d] = rnd<P>(r);
  }
}

// Zou-He velocity rows and inlet in storage arithmetic (fb/solvers/lbm.py:486-565).
template <class P>
__device__ __forceinline__ void unshift(typename P::C g[9]) {
  if constexpr (P::shift) {
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = g[d] + (typename P::C)D2Q9::w[d];
  }
}
template <class P>
__device__ __forceinline__ void reshift(typename P::C g[9]) {
  if constexpr (P::shift) {
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = g[d] - (typename P::C)D2Q9::w[d];
  }
}

// Shifted forms run Zou-He on h = f - w directly: every unknown pairs with a known population of the same
// weight (w2 = w4, w5 = w7, w6 = w8, w1 = w3), so the weights cancel in the updates, and the weights of the
// populations in each density sum add up to exactly 1, so rho = (1 + sum of the known h) / (1 -+ u). No value
// of the full magnitude f ~ w is formed in FP32, which keeps the edge cells' rounding at the scale of h, like
// the interior's.
template <class S, class T>
__device__ __forceinline__ void bottom_bc_shifted(const LbmArgs<S>& a, T g[9]) {
  const T ux = (T)a.tux, uy = (T)a.tuy;
  const T rho = ((T)1 + (g[0] + g[1] + g[3] + (T)2 * (g[4] + g[7] + g[8]))) / (T)(1.0 - a.tuy);
  g[2] = g[4] + (T)(2.0 / 3.0) * rho * uy;
  g[5] = g[7] + (T)0.5 * (g[3] - g[1]) + (T)(1.0 / 6.0) * rho * uy + (T)0.5 * rho * ux;
  g[6] = g[8] + (T)0.5 * (g[1] - g[3]) + (T)(1.0 / 6.0) * rho * uy - (T)0.5 * rho * ux;
}
template <class S, class T>
__device__ __forceinline__ void top_bc_shifted(const LbmArgs<S>& a, T g[9]) {
  const T ux = (T)a.tux, uy = (T)a.tuy;
  const T rho = ((T)1 + (g[0] + g[1] + g[3] + (T)2 * (g[2] + g[5] + g[6]))) / (T)(1.0 + a.tuy);
  g[4] = g[2] - (T)(2.0 / 3.0) * rho * uy;
  g[7] = g[5] + (T)0.5 * (g[1] - g[3]) - (T)(1.0 / 6.0) * rho * uy - (T)0.5 * rho * ux;
  g[8] = g[6] + (T)0.5 * (g[3] - g[1]) - (T)(1.0 / 6.0) * rho * uy + (T)0.5 * rho * ux;
}
template <class S, class T>
__device__ __forceinline__ void left_bc_shifted(const LbmArgs<S>& a, T g[9]) {
  const T ux = (T)a.tux, uy = (T)a.tuy;
  const T rho = ((T)1 + (g[0] + g[2] + g[4] + (T)2 * (g[3] + g[6] + g[7]))) / (T)(1.0 - a.tux);
  g[1] = g[3] + (T)(2.0 / 3.0) * rho * ux;
  g[5] = g[7] + (T)0.5 * (g[4] - g[2]) + (T)(1.0 / 6.0) * rho * ux + (T)0.5 * rho * uy;
  g[8] = g[6] + (T)0.5 * (g[2] - g[4]) + (T)(1.0 / 6.0) * rho * ux - (T)0.5 * rho * uy;
}

template <class P>
__device__ __forceinline__ void bottom_bc(const LbmArgs<typename P::S>& a, typename P::C g[9]) {
  if constexpr (P::shift) {
    bottom_bc_shifted(a, g);
    return;
  }
  using B = typename P::B;
  unshift<P>(g);
  B b[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) b[d] = (B)g[d];
  B ux = (B)a.tux, uy = (B)a.tuy;
  B rho = (b[0] + b[1] + b[3] + (B)2.0 * (b[4] + b[7] + b[8])) / (B)(1.0 - a.tuy);
  b[2] = b[4] + (B)(2.0 / 3.0) * rho * uy;
  b[5] = b[7] + (B)0.5 * (b[3] - b[1]) + (B)(1.0 / 6.0) * rho * uy + (B)0.5 * rho * ux;
  b[6] = b[8] + (B)0.5 * (b[1] - b[3]) + (B)(1.0 / 6.0) * rho * uy - (B)0.5 * rho * ux;
  g[2] = b[2];
  g[5] = b[5];
  g[6] = b[6];
  reshift<P>(g);
}

template <class P>
__device__ __forceinline__ void top_bc(const LbmArgs<typename P::S>& a, typename P::C g[9]) {
  if constexpr (P::shift) {
    top_bc_shifted(a, g);
    return;
  }
  using B = typename P::B;
  unshift<P>(g);
  B b[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) b[d] = (B)g[d];
  B ux = (B)a.tux, uy = (B)a.tuy;
  B rho = (b[0] + b[1] + b[3] + (B)2.0 * (b[2] + b[5] + b[6])) / (B)(1.0 + a.tuy);
  b[4] = b[2] - (B)(2.0 / 3.0) * rho * uy;
  b[7] = b[5] + (B)0.5 * (b[1] - b[3]) - (B)(1.0 / 6.0) * rho * uy - (B)0.5 * rho * ux;
  b[8] = b[6] + (B)0.5 * (b[3] - b[1]) - (B)(1.0 / 6.0) * rho * uy + (B)0.5 * rho * ux;
  g[4] = b[4];
  g[7] = b[7];
  g[8] = b[8];
  reshift<P>(g);
}

template <class P>
__device__ __forceinline__ void left_bc(const LbmArgs<typename P::S>& a, typename P::C g[9]) {
  if constexpr (P::shift) {
    left_bc_shifted(a, g);
    return;
  }
  using B = typename P::B;
  unshift<P>(g);
  B b[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) b[d] = (B)g[d];
  B ux = (B)a.tux, uy = (B)a.tuy;
  B rho = (b[0] + b[2] + b[4] + (B)2.0 * (b[3] + b[6] + b[7])) / (B)(1.0 - a.tux);
  b[1] = b[3] + (B)(2.0 / 3.0) * rho * ux;
  b[5] = b[7] + (B)0.5 * (b[4] - b[2]) + (B)(1.0 / 6.0) * rho * ux + (B)0.5 * rho * uy;
  b[8] = b[6] + (B)0.5 * (b[2] - b[4]) + (B)(1.0 / 6.0) * rho * ux - (B)0.5 * rho * uy;
  g[1] = b[1];
  g[5] = b[5];
  g[8] = b[8];
  reshift<P>(g);
}

// The cell whose pull an edge cell's boundary treatment starts from (fb/solvers/lbm.py:612-640):
// the cell itself; for the rows of channel walls, the row inside (mirrored); for the outlet
// column, the column before it (with the same row rule).
__device__ __forceinline__ void edge_source(int nx, int ny, int periodic_x, int periodic_y, int channel, int x,
                                            int y, int& px, int& py, bool& mirror) {
  px = (!periodic_x && x == nx - 1) ? nx - 2 : x;
  mirror = channel && !periodic_y && (y == 0 || y == ny - 1);
  py = mirror ? (y == 0 ? 1 : ny - 2) : y;
}

// Boundary treatment of edge cell (x,y) (fb/solvers/lbm.py:612-645) from h, the pull of its
// source cell (edge_source), and for the outlet column its history `hist`: mirrored wall rows or
// Zou-He rows, then the convective outlet (whose new history goes out through src.put_hist) or the
// Zou-He inlet, then the corner equilibria.
template <class P, class Src>
__device__ __forceinline__ void edge_fixup(const LbmArgs<typename P::S>& a, const Src& src, int x, int y,
                                           bool mirror, const typename P::C h[9], const typename P::S hist[9],
                                           typename P::C g[9]) {
  using S = typename P::S;
  using C = typename P::C;
  using B = typename P::B;
  if (mirror) {
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = h[D2Q9::opp[d]];
  } else {
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = h[d];
    if (!a.periodic_y && (y == 0 || y == a.ny - 1)) {
      if (y == 0) bottom_bc<P>(a, g);
      else top_bc<P>(a, g);
    }
  }
  if (!a.periodic_x && x == a.nx - 1) {
    const B u = (B)a.u_lattice;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      B prev = (B)ld<S, C>(&hist[d]);
      B v = prev + u * ((B)g[d] - prev);
      g[d] = (C)v;
      S sv;
      st<S, C>(&sv, (C)v);
      src.put_hist(y, d, sv);
    }
  } else if (!a.periodic_x && x == 0) {
    left_bc<P>(a, g);
  }
  if (!a.periodic_x && !a.periodic_y && (x == 0 || x == a.nx - 1) && (y == 0 || y == a.ny - 1)) {
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = (C)(a.feq_b[d] - (P::shift ? D2Q9::w[d] : 0.0));
  }
}

// Final pre-collision value of edge cell (x,y) for this step, before the sponge.
template <class P, class Src>
__device__ __forceinline__ void boundary_cell(const LbmArgs<typename P::S>& a, const Src& src, int x, int y,
                                              typename P::C g[9]) {
  using S = typename P::S;
  using C = typename P::C;
  int px, py;
  bool mirror;
  edge_source(a.nx, a.ny, a.periodic_x, a.periodic_y, a.channel, x, y, px, py, mirror);
  S hist[9];
  const bool outlet = !a.periodic_x && x == a.nx - 1;
#pragma unroll
  for (int d = 0; d < 9; ++d) hist[d] = outlet ? src.hist_at(y, d) : (S)0.0f;
  C h[9];
  pull_fast<P>(a, src, px, py, h);
  edge_fixup<P>(a, src, x, y, mirror, h, hist, g);
}

template <class P>
__device__ __forceinline__ void apply_sponge(const LbmArgs<typename P::S>& a, double sy, double sx,
                                             typename P::C g[9]) {
  using C = typename P::C;
  double s = fmax(fmax(0.0, sy), sx);
  if (s == 0.0) return;
  if constexpr (P::shift) {
    const C sf = (C)s, rf = (C)(1.0 - s);
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = rf * g[d] + sf * (C)(a.feq_b[d] - D2Q9::w[d]);
  } else {
    double r = 1.0 - s;
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = rnd<P>((C)(r * (double)g[d] + s * a.feq_b[d]));
  }
}

// TRT collision (kern/lbm.py:189-236). Returns the collision density.
template <class P>
__device__ __forceinline__ typename P::C collide(const LbmArgs<typename P::S>& a, const typename P::C g[9],
                                                 typename P::C post[9]) {
  using C = typename P::C;
  if constexpr (!P::shift) {
    C rho = g[0];
#pragma unroll
    for (int d = 1; d < 9; ++d) rho = rho + g[d];
    C safe = fmax(rho, (C)1.0e-12);
    C ux = ((((g[1] - g[3]) + g[5]) - g[6]) - g[7]) + g[8];
    ux = ux / safe;
    C uy = ((((g[2] - g[4]) + g[5]) + g[6]) - g[7]) - g[8];
    uy = uy / safe;
    C s2 = ux * ux + uy * uy;
    C feq[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      C p = (C)D2Q9::cx[d] * ux + (C)D2Q9::cy[d] * uy;
      feq[d] = rho * (C)D2Q9::w[d] * (((C)1.0 + (C)3.0 * p + (C)4.5 * p * p) - (C)1.5 * s2);
    }
    const C even = (C)a.even, odd = (C)a.odd;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const int o = D2Q9::opp[d];
      post[d] = g[d] - even * (g[d] - feq[d]) - odd * (g[o] - feq[o]);
    }
    return rho;
  } else {
    return C(0);
  }
}

#pragma clang fp contract(fast)

// Shifted-population TRT collision in the form's compute type T (FP32 for the shifted forms): h = f - w,
// rho = 1 + sum(h), h_eq = w (drho + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u^2)).
template <class P>
__device__ __forceinline__ typename P::C collide_shifted(const LbmArgs<typename P::S>& a, const typename P::C h[9],
                                                         typename P::C post[9]) {
  using T = typename P::C;
  T drho = h[0];
#pragma unroll
  for (int d = 1; d < 9; ++d) drho += h[d];
  const T rho = (T)1 + drho;
  const T inv = (T)1 / fmax(rho, (T)1.0e-12);
  const T ux = (((((h[1] - h[3]) + h[5]) - h[6]) - h[7]) + h[8]) * inv;
  const T uy = (((((h[2] - h[4]) + h[5]) + h[6]) - h[7]) - h[8]) * inv;
  const T s2 = ux * ux + uy * uy;
  T heq[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    const T p = (T)D2Q9::cx[d] * ux + (T)D2Q9::cy[d] * uy;
    heq[d] = (T)D2Q9::w[d] * (drho + rho * (((T)3 * p + (T)4.5 * p * p) - (T)1.5 * s2));
  }
  const T even = (T)a.even, odd = (T)a.odd;
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    const int o = D2Q9::opp[d];
    post[d] = h[d] - even * (h[d] - heq[d]) - odd * (h[o] - heq[o]);
  }
  return drho;
}

#pragma clang fp contract(off)

// Postcondition reductions of the pre-collision state g (fb/solvers/lbm.py:730-793):
// min population, finiteness, fluid density excursion, max physical speed. Density uses
// NumPy's pairwise order over the 9 directions and the momentum a sequential sum, in the
// storage dtype.
template <class P>
__device__ __forceinline__ void cell_stats(const LbmArgs<typename P::S>& a, const typename P::C gin[9], bool fluid,
                                           double& mn, double& vmax, double& exc, long long& nf, long long& bad,
                                           long long& nfu) {
  using B = typename P::B;
  using C = typename P::C;
  B f[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    C v = gin[d];
    if constexpr (P::shift) v = v + (C)D2Q9::w[d];
    f[d] = (B)v;
    double dv = (double)f[d];
    if (!isfinite(dv)) nf += 1;
    mn = fmin(mn, dv);
  }
  B rho = ((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7])) + f[8];
  B mx = (B)0, my = (B)0;
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    mx = mx + f[d] * (B)D2Q9::cx[d];
    my = my + f[d] * (B)D2Q9::cy[d];
  }
  B den = rho > (B)1.0e-12 ? rho : (B)1.0e-12;
  if (!(rho == rho)) den = rho;
  B ux = (mx / den) * (B)a.vscale;
  B uy = (my / den) * (B)a.vscale;
  B sp = sqrt(ux * ux + uy * uy);
  if (!isfinite((double)ux) || !isfinite((double)uy)) nfu += 1;
  vmax = fmax(vmax, (double)sp);
  if (fluid) {
    double r = (double)rho;
    if (!isfinite(r) || r <= 0.0) bad += 1;
    else exc = fmax(exc, (double)fabs(rho - (B)1.0));
  }
}

// ------------------------------------------------------------------ the step kernel

// Edge cell e of the lattice: bottom row, top row, then the left and right columns without
// their corners.
__device__ __forceinline__ void edge_cell(int e, int nx, int ny, int& x, int& y) {
  if (e < nx) {
    x = e;
    y = 0;
  } else if (e < 2 * nx) {
    x = e - nx;
    y = ny - 1;
  } else {
    int k = e - 2 * nx;
    if (k < ny - 2) {
      x = 0;
      y = 1 + k;
    } else {
      x = nx - 1;
      y = 1 + (k - (ny - 2));
    }
  }
}

// The edge blocks: the bottom and top rows 256 cells to a block, then the left and right columns
// kEdgeColumnLanes cells to a wave (lanes 0..kEdgeColumnLanes-1). A column cell's loads and stores
// each touch a cache line of their own (the lanes sit a row apart), so a block full of them is bound
// by its CU's address and cache-line throughput; spread a few to a wave over many CUs, the column
// cells finish within the interior's time instead of setting the length of the step.
constexpr int kEdgeColumnLanes = 16;
__host__ __device__ inline int edge_row_blocks(int nx) { return (2 * nx + 255) / 256; }
__host__ __device__ inline int edge_blocks_for(int nx, int ny) {
  const int per_block = 4 * kEdgeColumnLanes;
  return edge_row_blocks(nx) + (2 * (ny - 2) + per_block - 1) / per_block;
}
// Edge cell of thread `t` in edge block `blk` (-1: the thread has none).
__device__ __forceinline__ int edge_index(int blk, int t, int nx, int ny) {
  const int rb = edge_row_blocks(nx);
  if (blk < rb) {
    const int e = blk * 256 + t;
    return e < 2 * nx ? e : -1;
  }
  const int lane = t & 63;
  if (lane >= kEdgeColumnLanes) return -1;
  const int k = ((blk - rb) * 4 + (t >> 6)) * kEdgeColumnLanes + lane;
  return k < 2 * (ny - 2) ? 2 * nx + k : -1;
}

// Where a lattice step writes f* and the collision density of the near-wall cells: the global
// planes, or the LDS tile of the first step of a fused pair.
template <class S>
struct PlaneSink {
  S* f;
  S* rho;
  int nx;
  size_t N;
  template <class C>
  __device__ __forceinline__ void put(int d, int x, int y, C v) const {
    st<S, C>(f + (size_t)d * N + (size_t)y * nx + x, v);
  }
  template <class C>
  __device__ __forceinline__ void put_rho(int x, int y, C v) const {
    st<S, C>(rho + (size_t)y * nx + x, v);
  }
};

template <class S, int W, int H>
struct TileSink {
  S* f;    // [9][H][W]
  S* rho;  // [H][W]
  int ox, oy;
  template <class C>
  __device__ __forceinline__ void put(int d, int x, int y, C v) const {
    st<S, C>(f + d * (W * H) + (y - oy) * W + (x - ox), v);
  }
  template <class C>
  __device__ __forceinline__ void put_rho(int x, int y, C v) const {
    st<S, C>(rho + (y - oy) * W + (x - ox), v);
  }
};

template <class S>
__device__ __forceinline__ PlaneSink<S> plane_sink(const LbmArgs<S>& a) {
  return PlaneSink<S>{a.fout, a.rho_out, a.nx, (size_t)a.nx * a.ny};
}

// One cell of a lattice step: pull (straight for interior cells whose neighbours are all fluid,
// the boundary treatment for edge cells, the precomputed links for flagged cells), sponge,
// uncovered-cell reset, optional pre-collision write with the postcondition statistics, TRT
// collision, write f*. `prime` collides the pre-collision state in place instead of pulling.
template <class P, class Src, class Sink>
__device__ __forceinline__ void step_cell(const LbmArgs<typename P::S>& a, const Src& src, const Sink& out, int x,
                                          int y, bool straight, const double* links, bool solid, bool prime,
                                          bool reset, bool pre, double& mn, double& vmax, double& exc, long long& nf,
                                          long long& bad, long long& nfu) {
  using S = typename P::S;
  using C = typename P::C;
  const size_t N = (size_t)a.nx * a.ny;
  const size_t i = (size_t)y * a.nx + x;
  // operands that do not depend on the pull are loaded first, so the cell costs one round
  const double spy = a.sponge_y[y], spx = a.sponge_x[x];
  const uint8_t fnext = a.flags_next[i];
  const uint8_t fcur = reset ? a.flags[i] : (uint8_t)0;
  C g[9];
  if (prime) {
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const S v = src.at(d, x, y);
      g[d] = ld<S, C>(&v);
    }
  } else {
    if (straight) {
#pragma unroll
      for (int d = 0; d < 9; ++d) {
        const S v = src.at_in(d, x - D2Q9::cx[d], y - D2Q9::cy[d]);
        g[d] = ld<S, C>(&v);
      }
    } else if (links) {
      pull_linked<P>(a, src, x, y, solid, links, g);
    } else {
      boundary_cell<P>(a, src, x, y, g);
    }
    apply_sponge<P>(a, spy, spx, g);
  }
  if (reset && (fcur & FL_SOLID) && !(fnext & FL_SOLID)) {
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = (C)(a.feq_u[d] - (P::shift ? D2Q9::w[d] : 0.0));
  }
  if (pre) {
#pragma unroll
    for (int d = 0; d < 9; ++d) st<S, C>(a.fpre + (size_t)d * N + i, g[d]);
    cell_stats<P>(a, g, !(fnext & FL_SOLID), mn, vmax, exc, nf, bad, nfu);
  }
  C post[9];
  C rho;
  if constexpr (P::shift) rho = collide_shifted<P>(a, g, post);
  else rho = collide<P>(a, g, post);
#pragma unroll
  for (int d = 0; d < 9; ++d) out.put(d, x, y, post[d]);
  if (fnext & FL_NEAR) out.put_rho(x, y, rho);
}

// Block reduction of the postcondition statistics: wave shuffles, then one atomic per wave.
__device__ __forceinline__ void put_step_stats(unsigned long long* red, double mn, double vmax, double exc,
                                               long long nf, long long bad, long long nfu) {
  mn = wave_min(mn);
  vmax = wave_max(vmax);
  exc = wave_max(exc);
  nf = wave_sum_ll(nf);
  bad = wave_sum_ll(bad);
  nfu = wave_sum_ll(nfu);
  if ((threadIdx.x & 63) == 0) {
    key_min(&red[0], dkey(mn));
    key_max(&red[1], dkey(vmax));
    key_max(&red[2], dkey(exc));
    if (nf) atomicAdd(&red[3], (unsigned long long)nf);
    if (bad) atomicAdd(&red[4], (unsigned long long)bad);
    if (nfu) atomicAdd(&red[5], (unsigned long long)nfu);
  }
}

// Block reduction of the postcondition statistics into the block's slots of `spart` (order-free maxima,
// minima and counts), or the per-wave atomics when the launch has no slots. Every thread of the block
// calls it.
__device__ __forceinline__ void put_block_stats(unsigned long long* spart, unsigned long long* red, unsigned blk,
                                                unsigned nblk, double mn, double vmax, double exc, long long nf,
                                                long long bad, long long nfu) {
  if (!spart) {
    put_step_stats(red, mn, vmax, exc, nf, bad, nfu);
    return;
  }
  __shared__ unsigned long long ws[4][6];
  mn = wave_min(mn);
  vmax = wave_max(vmax);
  exc = wave_max(exc);
  nf = wave_sum_ll(nf);
  bad = wave_sum_ll(bad);
  nfu = wave_sum_ll(nfu);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    ws[w][0] = dkey(mn);
    ws[w][1] = dkey(vmax);
    ws[w][2] = dkey(exc);
    ws[w][3] = (unsigned long long)nf;
    ws[w][4] = (unsigned long long)bad;
    ws[w][5] = (unsigned long long)nfu;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    const int nw = (int)(blockDim.x >> 6);
    unsigned long long v = ws[0][q];
    for (int k = 1; k < nw; ++k) {
      const unsigned long long u = ws[k][q];
      v = q == 0 ? (u < v ? u : v) : q < 3 ? (u > v ? u : v) : v + u;
    }
    spart[(size_t)q * nblk + blk] = v;
  }
}

// Folds the per-block statistics of one step (lbm_step_kernel with `spart`) into `red`.
__global__ void __launch_bounds__(256) lbm_stats_fold_kernel(const unsigned long long* __restrict__ spart, int nblk,
                                                             unsigned long long* red) {
  __shared__ unsigned long long ws[4][6];
  unsigned long long v[6] = {~0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
  for (int b = (int)threadIdx.x; b < nblk; b += 256) {
    const unsigned long long k0 = spart[b], k1 = spart[(size_t)nblk + b], k2 = spart[2 * (size_t)nblk + b];
    v[0] = k0 < v[0] ? k0 : v[0];
    v[1] = k1 > v[1] ? k1 : v[1];
    v[2] = k2 > v[2] ? k2 : v[2];
    v[3] += spart[3 * (size_t)nblk + b];
    v[4] += spart[4 * (size_t)nblk + b];
    v[5] += spart[5 * (size_t)nblk + b];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const unsigned long long u = __shfl_xor(v[q], o, 64);
      v[q] = q == 0 ? (u < v[q] ? u : v[q]) : q < 3 ? (u > v[q] ? u : v[q]) : v[q] + u;
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 6; ++q) ws[w][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      v[0] = ws[k][0] < v[0] ? ws[k][0] : v[0];
      v[1] = ws[k][1] > v[1] ? ws[k][1] : v[1];
      v[2] = ws[k][2] > v[2] ? ws[k][2] : v[2];
      v[3] += ws[k][3];
      v[4] += ws[k][4];
      v[5] += ws[k][5];
    }
    atomicMin(&red[0], v[0]);
    atomicMax(&red[1], v[1]);
    atomicMax(&red[2], v[2]);
    atomicAdd(&red[3], v[3]);
    atomicAdd(&red[4], v[4]);
    atomicAdd(&red[5], v[5]);
  }
}

// ROWS interior cells of column x from row y0 (rows past ny-2 and columns 0, nx-1 idle): the
// straight pull of every cell, its flags, next-step flags and sponge weights are loaded before any
// cell is computed, so the cells share one round of loads; flagged cells are left to the list
// blocks. step_cell's arithmetic for straight cells (whose own flags are clear, so no reset).
template <class P, int ROWS, bool NT, bool PLAIN>
__device__ __forceinline__ void interior_cells(const LbmArgs<typename P::S>& a, int x, int y0, double& mn,
                                               double& vmax, double& exc, long long& nf, long long& bad,
                                               long long& nfu) {
  using S = typename P::S;
  using C = typename P::C;
  const size_t N = (size_t)a.nx * a.ny;
  const bool xin = x > 0 && x < a.nx - 1;
  S v[ROWS][9];
  uint8_t fl[ROWS], fnext[ROWS];
  double spy[ROWS];
  const double spx = xin ? a.sponge_x[x] : 0.0;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const int y = y0 + r;
    fl[r] = 1;
    if (xin && y <= a.ny - 2) {
      const size_t i = (size_t)y * a.nx + x;
#pragma unroll
      for (int d = 0; d < 9; ++d)
        v[r][d] = ld_pop<NT>(&a.fin[(size_t)d * N + (size_t)(y - D2Q9::cy[d]) * a.nx + (x - D2Q9::cx[d])]);
      fl[r] = a.flags[i];
      fnext[r] = a.flags_next[i];
      spy[r] = a.sponge_y[y];
    }
  }
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    if (fl[r] != 0) continue;
    const int y = y0 + r;
    const size_t i = (size_t)y * a.nx + x;
    C g[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = ld<S, C>(&v[r][d]);
    apply_sponge<P>(a, spy[r], spx, g);
    if (!PLAIN && a.write_pre) {
#pragma unroll
      for (int d = 0; d < 9; ++d) st<S, C>(a.fpre + (size_t)d * N + i, g[d]);
      cell_stats<P>(a, g, !(fnext[r] & FL_SOLID), mn, vmax, exc, nf, bad, nfu);
    }
    C post[9];
    C rho;
    if constexpr (P::shift) rho = collide_shifted<P>(a, g, post);
    else rho = collide<P>(a, g, post);
#pragma unroll
    for (int d = 0; d < 9; ++d) st_pop<NT, S, C>(a.fout + (size_t)d * N + i, post[d]);
    if (fnext[r] & FL_NEAR) st<S, C>(a.rho_out + i, rho);
  }
}

// One lattice step. The edge cells and the flagged cells (foil interior and its neighbours)
// run in the first blocks of the grid, so their longer dependent chains start early and do not
// diverge the interior waves; interior blocks take the straight pull.
// Block `blk` of the `nblk` blocks that cover lattice `a` for one step.
// PLAIN: a step without priming, uncovered-cell reset or pre-collision write (every step of an
// advance but its last, when the geometry holds still); those paths then cost no registers.
template <class P, int ROWS, bool NT, bool PLAIN = false>
__device__ __forceinline__ void lbm_block(const LbmArgs<typename P::S>& a, unsigned blk, unsigned nblk) {
  double mn = INFINITY, vmax = 0.0, exc = 0.0;
  long long nf = 0, bad = 0, nfu = 0;
  const bool prime = PLAIN ? false : (a.prime != 0);
  const bool reset = PLAIN ? false : (a.do_reset != 0);
  const bool pre = PLAIN ? false : (a.write_pre != 0);
  const unsigned special = (unsigned)(a.edge_blocks + a.list_blocks);
  if (blk < (unsigned)a.edge_blocks) {
    const int e = edge_index((int)blk, (int)threadIdx.x, a.nx, a.ny);
    if (e >= 0) {
      int x, y;
      edge_cell(e, a.nx, a.ny, x, y);
      step_cell<P>(a, plane_src(a), plane_sink(a), x, y, false, nullptr, false, prime, reset, pre, mn,
                   vmax, exc, nf, bad, nfu);
    }
  } else if (blk < special) {
    const int count = *a.list_count;
    const int stride = a.list_blocks * (int)blockDim.x;
    for (int k = (int)((blk - a.edge_blocks) * blockDim.x + threadIdx.x); k < count; k += stride) {
      const int entry = a.list[k];
      const int c = entry & 0x7fffffff;
      step_cell<P>(a, plane_src(a), plane_sink(a), c % a.nx, c / a.nx, false, a.links + (size_t)k * 8, entry < 0,
                   prime, reset, pre, mn, vmax, exc, nf, bad, nfu);
    }
  } else {
    const unsigned bi = xcd_remap(blk - special, nblk - special);
    const unsigned gx = (unsigned)a.int_gx;
    const int x = (int)((bi % gx) * blockDim.x + threadIdx.x);
    const int y0 = 1 + ROWS * (int)(bi / gx);
    if (prime) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        const int y = y0 + r;
        if (x > 0 && x < a.nx - 1 && y <= a.ny - 2 && a.flags[(size_t)y * a.nx + x] == 0)
          step_cell<P>(a, plane_src(a), plane_sink(a), x, y, true, nullptr, false, prime, reset, pre,
                       mn, vmax, exc, nf, bad, nfu);
      }
    } else {
      interior_cells<P, ROWS, NT, PLAIN>(a, x, y0, mn, vmax, exc, nf, bad, nfu);
    }
  }
  if (pre) put_block_stats(a.spart, a.red, blk, nblk, mn, vmax, exc, nf, bad, nfu);
}

// Waves per SIMD the step kernel is compiled for. The edge and Bouzidi paths need 88 registers in
// FP32 and 98 in shifted FP16 (five waves), the interior path 52; held to 80 (six waves) those
// paths spill a few values, which pays for FP32 and shifted FP16 (measured at 1024 x 512: 9.31 ->
// 8.96 and 7.60 -> 7.11 us per step) but not for FP64 (13.8 -> 15.0) or shifted FP32 (8.83 -> 8.96),
// which keep the compiler's allocation. Registers change no result bit.
template <class P>
struct StepWaves {
  static constexpr int min = (std::is_same<P, Prec<FB_F32>>::value || std::is_same<P, Prec<FB_F16_MIXED>>::value) ? 6 : 1;
};

// One lattice step with ROWS interior rows per block (a kernel of its own per ROWS, so that each
// gets its own register allocation); NT streams the interior populations past the caches.
template <class P, int ROWS, bool NT, bool PLAIN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(StepWaves<P>::min, 8)))
lbm_step_kernel(LbmArgs<typename P::S> a) {
  lbm_block<P, ROWS, NT, PLAIN>(a, blockIdx.x, gridDim.x);
}

// Blocks of one lattice step.
template <class S>
__host__ __device__ inline int lbm_step_blocks(const LbmArgs<S>& a) {
  return a.edge_blocks + a.list_blocks + interior_blocks(a.nx, a.ny, a.int_rows);
}

// One step of many lattices in one launch (one interior row per block). Lattice k's blocks start
// at start[k], a multiple of the XCD count, so that its tiles map onto the XCDs as in a launch of
// its own; the padding blocks between lattices return at once.
template <class P>
__global__ void __launch_bounds__(256) lbm_batch_kernel(const LbmArgs<typename P::S>* __restrict__ args,
                                                        const int* __restrict__ start, int n) {
  const int b = (int)blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const LbmArgs<typename P::S>& a = args[lo];
  const int nblk = lbm_step_blocks(a);
  const int local = b - start[lo];
  if (a.skip || local >= nblk) return;
  lbm_block<P, 1, false>(a, (unsigned)local, (unsigned)nblk);
}

// Postcondition statistics of the cells a thread has stepped.
struct StepStats {
  double mn = INFINITY, vmax = 0.0, exc = 0.0;
  long long nf = 0, bad = 0, nfu = 0;
};

// p - c wrapped onto [0, n) for |c| <= 1 (wrapi for these arguments, without a division).
__device__ __forceinline__ int wrap1(int p, int c, int n) {
  int s = p - c;
  s += s < 0 ? n : 0;
  s -= s >= n ? n : 0;
  return s;
}

// A cell of a fused pair that is not near the wall (the geometry unchanged over the pair): an edge
// cell from the pull v of its source cell (and the outlet history), a solid cell from its own rest
// population v[0], or a fluid cell from its straight pull v. step_cell's arithmetic for such cells.
template <class P, class Src, class Sink>
__device__ __forceinline__ void plain_cell(const LbmArgs<typename P::S>& a, const Src& src, const Sink& out, int x,
                                           int y, const typename P::S v[9], bool edge, bool mirror,
                                           const typename P::S hist[9], bool solid, double spy, double spx, bool pre,
                                           StepStats& acc) {
  using S = typename P::S;
  using C = typename P::C;
  C g[9];
  if (edge) {
    C h[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) h[d] = ld<S, C>(&v[d]);
    edge_fixup<P>(a, src, x, y, mirror, h, hist, g);
  } else {
    g[0] = ld<S, C>(&v[0]);
#pragma unroll
    for (int d = 1; d < 9; ++d) g[d] = solid ? (P::shift ? -(C)D2Q9::w[d] : (C)0) : ld<S, C>(&v[d]);
  }
  apply_sponge<P>(a, spy, spx, g);
  if (pre) {
    const size_t N = (size_t)a.nx * a.ny;
    const size_t i = (size_t)y * a.nx + x;
#pragma unroll
    for (int d = 0; d < 9; ++d) st<S, C>(a.fpre + (size_t)d * N + i, g[d]);
    cell_stats<P>(a, g, !solid, acc.mn, acc.vmax, acc.exc, acc.nf, acc.bad, acc.nfu);
  }
  C post[9];
  if constexpr (P::shift) (void)collide_shifted<P>(a, g, post);
  else (void)collide<P>(a, g, post);
#pragma unroll
  for (int d = 0; d < 9; ++d) out.put(d, x, y, post[d]);
}

__device__ __forceinline__ void merge_stats(StepStats& a, const StepStats& b) {
  a.mn = fmin(a.mn, b.mn);
  a.vmax = fmax(a.vmax, b.vmax);
  a.exc = fmax(a.exc, b.exc);
  a.nf += b.nf;
  a.bad += b.bad;
  a.nfu += b.nfu;
}

// Appends the lanes with `on` set to the block's LDS list: the cell's tile index and its list
// position k in the flagged-cell list.
__device__ __forceinline__ void list_append(bool on, int code, int k, unsigned short* codes, int* ks, int* count) {
  const unsigned long long m = __ballot(on);
  if (m == 0) return;
  const int lane = (int)threadIdx.x & 63;
  int base = 0;
  if (lane == 0) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, 0, 64);
  if (on) {
    const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
    codes[slot] = (unsigned short)code;
    ks[slot] = k;
  }
}

// Two lattice steps in one launch, for a non-periodic lattice whose geometry stays the same over
// both steps and whose flagged cells keep off the three outermost rows and columns. A block owns a
// tile of 64 x TH cells. The first step runs over the tile and a one-cell halo, reading the global
// f* planes, and leaves its f*, near-wall collision densities and outlet history in LDS; the second
// step runs over the tile from LDS and writes the global planes (and the statistics, on the last
// step of an advance). Every cell takes the single-step kernel's arithmetic, so the results are the
// same bits, and each step's f* crosses HBM once per two steps. The halo cells are stepped by this
// block and by their own: (TH + 2) * 66 first-step cells per 64 * TH output cells.
// In each step, one tile row per wave, every lane issues its pull (the straight pull, or for an
// edge cell the wrapped pull of its source cell) together with its flags, so fluid, solid and edge
// cells cost one round of loads; the near-wall cells, which need their Bouzidi links, are listed in
// LDS with their list positions and then run densely.
template <int TH>
struct PairTile {
  static constexpr int W = 66, H = TH + 2, A = W * H;
};
constexpr int kPairWaves = 8;
template <class S>
constexpr int pair_rows() {
  return sizeof(S) == 8 ? 8 : 16;
}

template <class P, int TH>
__global__ void __launch_bounds__(64 * kPairWaves) __attribute__((amdgpu_waves_per_eu(4, 8)))
lbm_pair_kernel(LbmArgs<typename P::S> a) {
  using S = typename P::S;
  constexpr int W = PairTile<TH>::W, H = PairTile<TH>::H, A = PairTile<TH>::A;
  constexpr int R2 = (TH + kPairWaves - 1) / kPairWaves;  // second-step rows per wave
  __shared__ S lf[9 * A];
  __shared__ S lrho[A];
  __shared__ S lhist[H * 9];
  __shared__ unsigned short lcode[A];
  __shared__ int lk[A];
  __shared__ int nlist[2];
  const unsigned ntx = (unsigned)(a.nx + 63) / 64;
  const unsigned bi = xcd_remap(blockIdx.x, gridDim.x);
  const int X0 = (int)(bi % ntx) * 64, Y0 = (int)(bi / ntx) * TH;
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  const int nx = a.nx, ny = a.ny;
  const size_t N = (size_t)nx * ny;
  if (threadIdx.x < 2) nlist[threadIdx.x] = 0;
  // second-step operands that do not depend on the first step, loaded ahead
  uint8_t fl2[R2];
  int k2[R2];
  double spy2[R2];
  const int x2 = X0 + lane;
  const double spx2 = x2 < nx ? a.sponge_x[x2] : 0.0;
#pragma unroll
  for (int r = 0; r < R2; ++r) {
    const int y = Y0 + wv + r * kPairWaves;
    const bool in = wv + r * kPairWaves < TH && x2 < nx && y < ny;
    const bool edge = x2 == 0 || y == 0 || x2 == nx - 1 || y == ny - 1;
    fl2[r] = (in && !edge) ? a.flags[(size_t)y * nx + x2] : (uint8_t)0;
    spy2[r] = in ? a.sponge_y[y] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < R2; ++r) {
    const int y = Y0 + wv + r * kPairWaves;
    k2[r] = (fl2[r] & FL_NEAR) ? a.lidx[(size_t)y * nx + x2] : 0;
  }
  StepStats st;
  __syncthreads();
  {
    // first step: H rows of the 64 tile columns, then the 2 H halo-column cells
    const PlaneSrc<S> src{a.fin, a.rho_in, a.outlet_in, lhist, Y0 - 1, nx, N};
    const TileSink<S, W, H> out{lf, lrho, X0 - 1, Y0 - 1};
    constexpr int jobs = H + (2 * H + 63) / 64;
    for (int j = wv; j < jobs; j += kPairWaves) {
      int lx, ly;
      bool in = true;
      if (j < H) {
        lx = 1 + lane;
        ly = j;
      } else {
        const int h = (j - H) * 64 + lane;
        in = h < 2 * H;
        ly = h >> 1;
        lx = (h & 1) ? W - 1 : 0;
      }
      const int x = X0 - 1 + lx, y = Y0 - 1 + ly;
      in = in && x >= 0 && x < nx && y >= 0 && y < ny;
      const bool edge = x == 0 || y == 0 || x == nx - 1 || y == ny - 1;
      int px = x, py = y;
      bool mirror = false;
      if (edge) edge_source(nx, ny, 0, 0, a.channel, x, y, px, py, mirror);
      S v[9], hist[9];
      uint8_t fl = 0;
      double spy = 0.0, spx = 0.0;
      if (in) {
#pragma unroll
        for (int d = 0; d < 9; ++d)
          v[d] = a.fin[(size_t)d * N + (size_t)wrap1(py, D2Q9::cy[d], ny) * nx + wrap1(px, D2Q9::cx[d], nx)];
        const bool outlet = x == nx - 1;
#pragma unroll
        for (int d = 0; d < 9; ++d) hist[d] = outlet ? a.outlet_in[(size_t)y * 9 + d] : (S)0.0f;
        if (!edge) fl = a.flags[(size_t)y * nx + x];
        spy = a.sponge_y[y];
        spx = a.sponge_x[x];
      }
      const bool near = (fl & FL_NEAR) != 0;
      const int k = near ? a.lidx[(size_t)y * nx + x] : 0;
      if (in && !near) plain_cell<P>(a, src, out, x, y, v, edge, mirror, hist, (fl & FL_SOLID) != 0, spy, spx, false, st);
      list_append(in && near, ly * W + lx, k, lcode, lk, &nlist[0]);
    }
    __syncthreads();
    const int n = nlist[0];
    for (int m = (int)threadIdx.x; m < n; m += 64 * kPairWaves) {
      const int code = lcode[m];
      const int x = X0 - 1 + code % W, y = Y0 - 1 + code / W;
      step_cell<P>(a, src, out, x, y, false, a.links + (size_t)lk[m] * 8, false, false, false, false, st.mn, st.vmax,
                   st.exc, st.nf, st.bad, st.nfu);
    }
  }
  __syncthreads();
  {
    const TileSrc<S, W, H> src{lf, lrho, lhist, a.outlet_out, X0 - 1, Y0 - 1};
    const PlaneSink<S> out = plane_sink(a);
    const bool pre = a.write_pre != 0;
#pragma unroll
    for (int r = 0; r < R2; ++r) {
      const int ry = wv + r * kPairWaves;
      const int x = x2, y = Y0 + ry;
      const bool in = ry < TH && x < nx && y < ny;
      const bool edge = x == 0 || y == 0 || x == nx - 1 || y == ny - 1;
      const bool near = (fl2[r] & FL_NEAR) != 0;
      if (in && !near) {
        int px = x, py = y;
        bool mirror = false;
        if (edge) edge_source(nx, ny, 0, 0, a.channel, x, y, px, py, mirror);
        S v[9], hist[9];
#pragma unroll
        for (int d = 0; d < 9; ++d) {
          const int sx = wrap1(px, D2Q9::cx[d], nx), sy = wrap1(py, D2Q9::cy[d], ny);
          v[d] = edge ? src.at(d, sx, sy) : src.at_in(d, sx, sy);
        }
#pragma unroll
        for (int d = 0; d < 9; ++d) hist[d] = x == nx - 1 ? src.hist_at(y, d) : (S)0.0f;
        plain_cell<P>(a, src, out, x, y, v, edge, mirror, hist, (fl2[r] & FL_SOLID) != 0, spy2[r], spx2, pre, st);
      }
      list_append(in && near, (ry + 1) * W + lane + 1, k2[r], lcode, lk, &nlist[1]);
    }
    __syncthreads();
    const int n = nlist[1];
    for (int m = (int)threadIdx.x; m < n; m += 64 * kPairWaves) {
      const int code = lcode[m];
      const int x = X0 - 1 + code % W, y = Y0 - 1 + code / W;
      step_cell<P>(a, src, out, x, y, false, a.links + (size_t)lk[m] * 8, false, false, false, pre, st.mn, st.vmax,
                   st.exc, st.nf, st.bad, st.nfu);
    }
  }
  if (a.write_pre) put_step_stats(a.red, st.mn, st.vmax, st.exc, st.nf, st.bad, st.nfu);
}

// The flagged cells of rows 1..ny-2, columns 1..nx-2, listed in row-major order so that
// neighbouring lanes of the list blocks touch neighbouring cells. Three passes: count per
// row, scan over rows, fill. The sign bit of an entry marks a solid cell.
__device__ __forceinline__ bool listed(const uint8_t* flags, int nx, int ny, int x, int y) {
  return x >= 1 && x < nx - 1 && y >= 1 && y < ny - 1 && flags[(size_t)y * nx + x] != 0;
}

__global__ void __launch_bounds__(256) lbm_row_count_kernel(int nx, int ny, const uint8_t* __restrict__ flags,
                                                            int* __restrict__ row_count) {
  __shared__ int part[4];
  const int y = blockIdx.x;
  int c = 0;
  for (int x = threadIdx.x; x < nx; x += 256) c += listed(flags, nx, ny, x, y) ? 1 : 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) row_count[y] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of the row counts (one block); the total into *count.
__global__ void __launch_bounds__(256) lbm_row_scan_kernel(int ny, const int* __restrict__ row_count,
                                                           int* __restrict__ row_start, int* __restrict__ count) {
  __shared__ int wsum[4];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int base = 0; base < ny; base += 256) {
    const int y = base + (int)threadIdx.x;
    const int v = y < ny ? row_count[y] : 0;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int off = carry;
    for (int j = 0; j < w; ++j) off += wsum[j];
    if (y < ny) row_start[y] = off + incl - v;
    __syncthreads();
    if (threadIdx.x == 255) carry = off + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = carry;
}

__global__ void __launch_bounds__(256) lbm_row_fill_kernel(int nx, int ny, const uint8_t* __restrict__ flags,
                                                           const int* __restrict__ row_start,
                                                           int* __restrict__ list, int* __restrict__ lidx) {
  __shared__ int wsum[4];
  const int y = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int run = row_start[y];
  for (int base = 0; base < nx; base += 256) {
    const int x = base + (int)threadIdx.x;
    const bool on = x < nx && listed(flags, nx, ny, x, y);
    const unsigned long long mask = __ballot(on);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(mask);
    __syncthreads();
    int off = run;
    for (int j = 0; j < w; ++j) off += wsum[j];
    if (on) {
      const size_t i = (size_t)y * nx + x;
      list[off + before] = (int)i | ((flags[i] & FL_SOLID) ? (int)0x80000000u : 0);
      lidx[i] = off + before;
    }
    run += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
}

// Bouzidi fractions of the listed cells' solid links, with reflect()'s expression.
__global__ void __launch_bounds__(256) lbm_links_kernel(int nx, int ny, const uint8_t* __restrict__ flags,
                                                        const double* __restrict__ sdf, const int* __restrict__ list,
                                                        const int* __restrict__ count, double* __restrict__ links) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= *count) return;
  const int c = list[k] & 0x7fffffff;
  const int x = c % nx, y = c / nx;
  const bool fluid_near = !(flags[c] & FL_SOLID) && (flags[c] & FL_NEAR);
  const double sx = sdf[c];
  for (int d = 1; d < 9; ++d) {
    double q = 0.0;
    const int px = wrapi(x - D2Q9::cx[d], nx), py = wrapi(y - D2Q9::cy[d], ny);
    const size_t s = (size_t)py * nx + px;
    if (fluid_near && (flags[s] & FL_SOLID)) {
      double sn = sdf[s];
      double den = fmax(sx - sn, 1.0e-12);
      q = fmin(fmax(sx / den, 0.05), 1.0);
    }
    links[(size_t)k * 8 + d - 1] = q;
  }
}

// Near-wall flags: fluid cells with a solid cell among their 8 periodic neighbours.
__global__ void __launch_bounds__(256) lbm_flags_kernel(int nx, int ny, const uint8_t* __restrict__ solid,
                                                        uint8_t* __restrict__ flags) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= nx) return;
  size_t i = (size_t)y * nx + x;
  uint8_t f = solid[i] ? FL_SOLID : 0;
  if (!f) {
    for (int d = 1; d < 9; ++d) {
      int sx = wrapi(x + D2Q9::cx[d], nx), sy = wrapi(y + D2Q9::cy[d], ny);
      if (solid[(size_t)sy * nx + sx]) f |= FL_NEAR;
    }
  }
  flags[i] = f;
}

// Rescale (fb/solvers/lbm.py:357-368) on the pre-collision field, storage arithmetic.
template <class P>
__global__ void __launch_bounds__(256) lbm_rescale_kernel(int nx, int ny, const typename P::S* __restrict__ fin,
                                                          typename P::S* __restrict__ fout, double ratio,
                                                          typename P::S* __restrict__ outlet) {
  using S = typename P::S;
  using C = typename P::C;
  using B = typename P::B;
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= nx) return;
  const size_t N = (size_t)nx * ny;
  const size_t i = (size_t)y * nx + x;
  B f[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    C v = ld<S, C>(fin + (size_t)d * N + i);
    if constexpr (P::shift) v = v + (C)D2Q9::w[d];
    f[d] = (B)v;
  }
  B rho = ((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7])) + f[8];
  B mx = (B)0, my = (B)0;
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    mx = mx + f[d] * (B)D2Q9::cx[d];
    my = my + f[d] * (B)D2Q9::cy[d];
  }
  B den = rho > (B)1.0e-12 ? rho : (B)1.0e-12;
  B ux = mx / den, uy = my / den;
  B r = (B)ratio;
  B nux = ux * r, nuy = uy * r;
  B s2o = ux * ux + uy * uy;
  B s2n = nux * nux + nuy * nuy;
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    B cx = (B)D2Q9::cx[d], cy = (B)D2Q9::cy[d], w = (B)D2Q9::w[d];
    B cuo = cx * ux + cy * uy;
    B cun = cx * nux + cy * nuy;
    B eo = rho * w * ((B)1.0 + (B)3.0 * cuo + (B)4.5 * cuo * cuo - (B)1.5 * s2o);
    B en = rho * w * ((B)1.0 + (B)3.0 * cun + (B)4.5 * cun * cun - (B)1.5 * s2n);
    B v = en + r * (f[d] - eo);
    C out = (C)v;
    if constexpr (P::shift) out = out - (C)D2Q9::w[d];
    st<S, C>(fout + (size_t)d * N + i, out);
    if (x == nx - 1) st<S, C>(outlet + (size_t)y * 9 + d, out);
  }
}

// Density and physical velocity for export/sampling (fb/solvers/lbm.py:447-461).
template <class P>
__global__ void __launch_bounds__(256) lbm_macro_kernel(int nx, int ny, const typename P::S* __restrict__ fin,
                                                        double vscale, typename P::B* __restrict__ rho_out,
                                                        typename P::B* __restrict__ vel_out) {
  using S = typename P::S;
  using C = typename P::C;
  using B = typename P::B;
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= nx) return;
  const size_t N = (size_t)nx * ny;
  const size_t i = (size_t)y * nx + x;
  if constexpr (P::shift) {
    // moments of h = f - w in FP64 (the weights contribute 1 to the density and nothing to the momentum)
    double h[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) h[d] = (double)ld<S, C>(fin + (size_t)d * N + i);
    const double dr = ((h[0] + h[1]) + (h[2] + h[3])) + ((h[4] + h[5]) + (h[6] + h[7])) + h[8];
    const double jx = ((((h[1] - h[3]) + h[5]) - h[6]) - h[7]) + h[8];
    const double jy = ((((h[2] - h[4]) + h[5]) + h[6]) - h[7]) - h[8];
    const double rho = 1.0 + dr;
    const double den = rho > 1.0e-12 ? rho : 1.0e-12;
    rho_out[i] = (B)rho;
    vel_out[2 * i] = (B)(jx / den * vscale);
    vel_out[2 * i + 1] = (B)(jy / den * vscale);
    return;
  }
  B f[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) f[d] = (B)ld<S, C>(fin + (size_t)d * N + i);
  B rho = ((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7])) + f[8];
  B mx = (B)0, my = (B)0;
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    mx = mx + f[d] * (B)D2Q9::cx[d];
    my = my + f[d] * (B)D2Q9::cy[d];
  }
  B den = rho > (B)1.0e-12 ? rho : (B)1.0e-12;
  rho_out[i] = rho;
  vel_out[2 * i] = (mx / den) * (B)vscale;
  vel_out[2 * i + 1] = (my / den) * (B)vscale;
}

// Pose change outside an advance: uncovered cells (solid before, fluid after) take the
// equilibrium feq_u (fb/solvers/lbm.py:463-484); other cells are copied.
template <class P>
__global__ void __launch_bounds__(256) lbm_uncover_kernel(int nx, int ny, const uint8_t* __restrict__ old_flags,
                                                          const uint8_t* __restrict__ new_flags,
                                                          const typename P::S* __restrict__ fin,
                                                          typename P::S* __restrict__ fout, LbmFeq feq_u) {
  using S = typename P::S;
  using C = typename P::C;
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= nx) return;
  const size_t N = (size_t)nx * ny;
  const size_t i = (size_t)y * nx + x;
  const bool unc = (old_flags[i] & FL_SOLID) && !(new_flags[i] & FL_SOLID);
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    if (unc) {
      C v = (C)feq_u.v[d];
      if constexpr (P::shift) v = v - (C)D2Q9::w[d];
      st<S, C>(fout + (size_t)d * N + i, v);
    } else {
      fout[(size_t)d * N + i] = fin[(size_t)d * N + i];
    }
  }
}

// f <- f_eq(rho, u) in storage arithmetic (fb/solvers/lbm.py:428-445); outlet <- f[:, -1].
template <class P>
__global__ void __launch_bounds__(256) lbm_equilibrium_kernel(int nx, int ny, const typename P::B* __restrict__ rho_in,
                                                              const typename P::B* __restrict__ vel_in,
                                                              typename P::S* __restrict__ fout,
                                                              typename P::S* __restrict__ outlet) {
  using S = typename P::S;
  using C = typename P::C;
  using B = typename P::B;
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= nx) return;
  const size_t N = (size_t)nx * ny;
  const size_t i = (size_t)y * nx + x;
  B rho = rho_in[i];
  B ux = vel_in[2 * i], uy = vel_in[2 * i + 1];
  if constexpr (P::shift) {
    // h = f_eq - w = w ((rho - 1) + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u^2)), formed in FP64 and rounded once to the
    // storage type: the stored deviation is then exact to its own rounding rather than carrying the rounding
    // of an FP32 value of size w.
    const double r = (double)rho, vx = (double)ux, vy = (double)uy;
    const double q2 = vx * vx + vy * vy;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const double cu = (double)D2Q9::cx[d] * vx + (double)D2Q9::cy[d] * vy;
      const double h = D2Q9::w[d] * ((r - 1.0) + r * ((3.0 * cu + 4.5 * cu * cu) - 1.5 * q2));
      st<S, C>(fout + (size_t)d * N + i, (C)h);
      if (x == nx - 1) st<S, C>(outlet + (size_t)y * 9 + d, (C)h);
    }
    return;
  }
  B s2 = ux * ux + uy * uy;
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    B cu = (B)D2Q9::cx[d] * ux + (B)D2Q9::cy[d] * uy;
    B e = rho * (B)D2Q9::w[d] * ((B)1.0 + (B)3.0 * cu + (B)4.5 * cu * cu - (B)1.5 * s2);
    C out = (C)e;
    st<S, C>(fout + (size_t)d * N + i, out);
    if (x == nx - 1) st<S, C>(outlet + (size_t)y * 9 + d, out);
  }
}

// Host AoS (ny, nx, 9) <-> device SoA planes, with optional shift.
template <class P>
__global__ void __launch_bounds__(256) lbm_aos_to_soa(int nx, int ny, const typename P::B* __restrict__ aos,
                                                      typename P::S* __restrict__ soa) {
  using C = typename P::C;
  size_t N = (size_t)nx * ny;
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N * 9) return;
  size_t i = k / 9;
  int d = (int)(k % 9);
  C v = (C)aos[k];
  if constexpr (P::shift) v = v - (C)D2Q9::w[d];
  st<typename P::S, C>(soa + (size_t)d * N + i, v);
}
template <class P>
__global__ void __launch_bounds__(256) lbm_soa_to_aos(int nx, int ny, const typename P::S* __restrict__ soa,
                                                      typename P::B* __restrict__ aos) {
  using C = typename P::C;
  size_t N = (size_t)nx * ny;
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N * 9) return;
  size_t i = k / 9;
  int d = (int)(k % 9);
  if constexpr (P::shift) {
    aos[k] = (typename P::B)((double)ld<typename P::S, C>(soa + (size_t)d * N + i) + D2Q9::w[d]);
  } else {
    aos[k] = (typename P::B)ld<typename P::S, C>(soa + (size_t)d * N + i);
  }
}
template <class P>
__global__ void lbm_outlet_convert(int ny, const typename P::B* __restrict__ in_b, typename P::S* __restrict__ out_s,
                                   const typename P::S* __restrict__ in_s, typename P::B* __restrict__ out_b) {
  using C = typename P::C;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ny * 9) return;
  int d = k % 9;
  if (in_b) {
    C v = (C)in_b[k];
    if constexpr (P::shift) v = v - (C)D2Q9::w[d];
    st<typename P::S, C>(out_s + k, v);
  } else {
    C v = ld<typename P::S, C>(in_s + k);
    if constexpr (P::shift) v = v + (C)D2Q9::w[d];
    out_b[k] = (typename P::B)v;
  }
}

// ------------------------------------------------------------------ host handle

// Blocks that stride over the flagged-cell list of a lattice step.
constexpr int kListBlocks = 32;

struct Geometry {
  double* sdf = nullptr;
  uint8_t* solid = nullptr;
  uint8_t* flags = nullptr;
  int* list = nullptr;        // flagged cells of rows 1..ny-2 (device-built, unordered)
  int* list_count = nullptr;
  int* rows = nullptr;        // per-row counts, then per-row starts, of the list
  double* links = nullptr;    // 8 Bouzidi fractions per list entry
  int* lidx = nullptr;        // list position of each listed cell
  double angle = 0.0;
  bool valid = false;
};

struct Lattice {
  Ctx* ctx = nullptr;
  fb_domain dom{};
  fb_foil foil{};
  int prec = FB_F64;
  int channel = 0;
  size_t N = 0;
  size_t ssize = 8;     // bytes per stored value
  size_t bsize = 8;     // bytes per host value
  void* F[3] = {};      // post-collision planes
  void* R[3] = {};      // collision density (near-wall cells)
  void* O[3] = {};      // outlet history (ny*9)
  void* PRE[3] = {};    // pre-collision planes
  Geometry G[3];
  double* sponge_y = nullptr;
  double* sponge_x = nullptr;
  unsigned long long* red = nullptr;
  unsigned long long* red_host = nullptr;
  unsigned long long* spart = nullptr;  // per-block statistics slots of a step that writes the pre-collision state
  int spart_blocks = 0;                  // blocks the slots hold
  void* host_stage = nullptr;
  // current indices
  int f = 0, o = 0, pre = 0, g = 0;
  bool fstar_valid = false;  // F[f] == collide(PRE[pre]) for geometry G[g]
  // checkpoint
  int cf = 0, co = 0, cpre = 0, cg = 0;
  bool cvalid = false, cfstar = false;
};

// Blocks striding over the flagged-cell list (tuned per lattice shape; each flagged cell is
// stepped by the same arithmetic whichever block takes it).
static int list_blocks(const Lattice& L) {
  const int t = tune_get(TUNE_LBM_LIST_BLOCKS, L.dom.nx, L.dom.ny, L.prec, 1);
  return (t >= 1 && t <= 1024) ? t : kListBlocks;
}

// Interior rows per thread of the single-step kernel: two cells per thread give each wave more
// loads in flight, which pays on large lattices; small lattices want the extra waves. Tuned per
// shape (1 or 2), default 2 from 2 Mi cells; the choice changes no result.
static int int_rows(const Lattice& L) {
  const int t = tune_get(TUNE_LBM_ROWS, L.dom.nx, L.dom.ny, L.prec, 1);
  if (t == 1 || t == 2) return t;
  return L.N >= ((size_t)1 << 21) ? 2 : 1;
}

// Nontemporal population traffic in the single-step kernel: on by default once the two population
// copies outgrow the 256 MB Infinity Cache (measured at 4096 x 2048 FP32, 140 -> 126 us per step), off
// below it, where the next step finds the populations in the cache; tuned per shape, no result changes.
static bool stream_pops(const Lattice& L) {
  const int t = tune_get(TUNE_LBM_NT, L.dom.nx, L.dom.ny, L.prec, 1);
  if (t == 0 || t == 1) return t == 1;
  return 2 * 9 * L.N * L.ssize > ((size_t)192 << 20);
}

static int other3(int a, int b) {
  for (int k = 0; k < 3; ++k)
    if (k != a && k != b) return k;
  return 0;
}

template <class S>
static LbmArgs<S> step_args(Lattice& L, const fb_lbm_params& p, int fin_idx, bool prime, int fout_idx, int o_in,
                            int o_out, int g_cur, int g_next, bool write_pre, int pre_out, bool do_reset) {
  LbmArgs<S> a{};
  a.nx = L.dom.nx;
  a.ny = L.dom.ny;
  a.fin = prime ? (const S*)L.PRE[L.pre] : (const S*)L.F[fin_idx];
  a.fout = (S*)L.F[fout_idx];
  a.fpre = (S*)L.PRE[pre_out];
  a.rho_in = (const S*)L.R[fin_idx];
  a.rho_out = (S*)L.R[fout_idx];
  a.outlet_in = (const S*)L.O[o_in];
  a.outlet_out = (S*)L.O[o_out];
  a.flags = L.G[g_cur].flags;
  a.flags_next = L.G[g_next].flags;
  a.sdf = L.G[g_cur].sdf;
  a.sponge_y = L.sponge_y;
  a.sponge_x = L.sponge_x;
  a.x0 = L.dom.x0;
  a.y0 = L.dom.y0;
  a.dx = (L.dom.x1 - L.dom.x0) / L.dom.nx;
  a.dy = (L.dom.y1 - L.dom.y0) / L.dom.ny;
  a.pivot_x = L.foil.pivot_x;
  a.pivot_y = L.foil.pivot_y;
  a.wall_omega = p.wall_omega_rad;
  a.wall_scale = p.wall_scale;
  a.even = 0.5 * (p.omega_plus + p.omega_minus);
  a.odd = 0.5 * (p.omega_plus - p.omega_minus);
  a.u_lattice = p.u_lattice;
  a.tux = p.target_ux;
  a.tuy = p.target_uy;
  a.vscale = p.velocity_scale;
  for (int d = 0; d < 9; ++d) {
    a.feq_b[d] = p.feq_boundary[d];
    a.feq_u[d] = p.feq_uncovered[d];
  }
  a.periodic_x = L.dom.periodic_x;
  a.periodic_y = L.dom.periodic_y;
  a.channel = L.channel;
  a.write_pre = write_pre ? 1 : 0;
  a.do_reset = do_reset ? 1 : 0;
  a.prime = prime ? 1 : 0;
  a.red = L.red;
  a.int_gx = (L.dom.nx + 255) / 256;
  a.int_rows = int_rows(L);
  a.edge_blocks = edge_blocks_for(L.dom.nx, L.dom.ny);
  a.list_blocks = list_blocks(L);
  a.list = L.G[g_cur].list;
  a.list_count = L.G[g_cur].list_count;
  a.links = L.G[g_cur].links;
  a.lidx = L.G[g_cur].lidx;
  a.skip = 0;
  a.spart = nullptr;
  return a;
}

template <class P>
static int launch_step(Lattice& L, const fb_lbm_params& p, int fin_idx, bool prime, int fout_idx, int o_in, int o_out,
                       int g_cur, int g_next, bool write_pre, int pre_out, bool do_reset) {
  using S = typename P::S;
  LbmArgs<S> a =
      step_args<S>(L, p, fin_idx, prime, fout_idx, o_in, o_out, g_cur, g_next, write_pre, pre_out, do_reset);
  const dim3 grid(lbm_step_blocks(a));
  // the statistics of a step that writes the pre-collision state go through per-block slots and one fold
  // (per-wave atomics on the six keys serialise on large grids)
  const bool slots = write_pre && !prime && (int)grid.x <= L.spart_blocks;
  if (slots) a.spart = L.spart;
  const bool plain = !prime && !write_pre && !do_reset;
  const bool nt = stream_pops(L);
  auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, dim3(256), 0, L.ctx->stream, a); };
  if (plain) {
    if (a.int_rows == 2) nt ? go(lbm_step_kernel<P, 2, true, true>) : go(lbm_step_kernel<P, 2, false, true>);
    else nt ? go(lbm_step_kernel<P, 1, true, true>) : go(lbm_step_kernel<P, 1, false, true>);
  } else {
    if (a.int_rows == 2) nt ? go(lbm_step_kernel<P, 2, true, false>) : go(lbm_step_kernel<P, 2, false, false>);
    else nt ? go(lbm_step_kernel<P, 1, true, false>) : go(lbm_step_kernel<P, 1, false, false>);
  }
  if (slots)
    hipLaunchKernelGGL(lbm_stats_fold_kernel, dim3(1), dim3(256), 0, L.ctx->stream, (const unsigned long long*)L.spart,
                       (int)grid.x, L.red);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

// Whether two steps of this lattice may run as one fused launch (lbm_pair_kernel): a
// non-periodic lattice whose flagged cells keep off the three outermost rows and columns, and
// whose last tile column and row hold at least two cells. A NACA four-digit section lies within
// 0.96 chord of its pivot (quarter chord; |y| <= (m + t/2) c), so a pivot more than a chord plus
// eight cells inside every edge suffices. Fusion halves the HBM traffic per step, but a block
// then runs three dependent rounds of loads and the edge and foil tiles run longer than the rest,
// so it is chosen per shape by the autotuner (tuning value 1) and off otherwise; the choice
// changes no result.
template <class S>
static bool pair_ok(const Lattice& L) {
  if (tune_get(TUNE_LBM_FUSE, L.dom.nx, L.dom.ny, L.prec, 1) != 1) return false;
  if (L.dom.periodic_x || L.dom.periodic_y) return false;
  if (L.dom.nx % 64 == 1 || L.dom.ny % pair_rows<S>() == 1) return false;
  const double dx = (L.dom.x1 - L.dom.x0) / L.dom.nx, dy = (L.dom.y1 - L.dom.y0) / L.dom.ny;
  const double r = L.foil.chord;
  return L.foil.pivot_x - r - 8.0 * dx > L.dom.x0 && L.foil.pivot_x + r + 8.0 * dx < L.dom.x1 &&
         L.foil.pivot_y - r - 8.0 * dy > L.dom.y0 && L.foil.pivot_y + r + 8.0 * dy < L.dom.y1;
}

// Steps k and k+1 of an advance as one launch: f* from F[fin_idx] to F[fout_idx] (the first
// step's f* stays in LDS), outlet history from O[o_in] to O[o_out].
template <class P>
static int launch_pair(Lattice& L, const fb_lbm_params& p, int fin_idx, int fout_idx, int o_in, int o_out, int g,
                       bool write_pre, int pre_out) {
  using S = typename P::S;
  constexpr int th = pair_rows<S>();
  const LbmArgs<S> a = step_args<S>(L, p, fin_idx, false, fout_idx, o_in, o_out, g, g, write_pre, pre_out, false);
  const unsigned blocks = (unsigned)((L.dom.nx + 63) / 64) * (unsigned)((L.dom.ny + th - 1) / th);
  hipLaunchKernelGGL((lbm_pair_kernel<P, th>), dim3(blocks), dim3(64 * kPairWaves), 0, L.ctx->stream, a);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

static int compute_geometry(Lattice& L, int gi, const fb_pose& pose) {
  Geometry& G = L.G[gi];
  int rc = launch_cell_sdf(L.ctx->stream, make_cell_grid(L.dom), make_naca(L.foil, pose), G.sdf, G.solid);
  if (rc) return rc;
  dim3 grid((L.dom.nx + 255) / 256, L.dom.ny);
  hipLaunchKernelGGL(lbm_flags_kernel, grid, dim3(256), 0, L.ctx->stream, L.dom.nx, L.dom.ny, G.solid, G.flags);
  hipLaunchKernelGGL(lbm_row_count_kernel, dim3(L.dom.ny), dim3(256), 0, L.ctx->stream, L.dom.nx, L.dom.ny,
                     G.flags, G.rows);
  hipLaunchKernelGGL(lbm_row_scan_kernel, dim3(1), dim3(256), 0, L.ctx->stream, L.dom.ny, G.rows,
                     G.rows + L.dom.ny, G.list_count);
  hipLaunchKernelGGL(lbm_row_fill_kernel, dim3(L.dom.ny), dim3(256), 0, L.ctx->stream, L.dom.nx, L.dom.ny, G.flags,
                     G.rows + L.dom.ny, G.list, G.lidx);
  hipLaunchKernelGGL(lbm_links_kernel, dim3((unsigned)((L.N + 255) / 256)), dim3(256), 0, L.ctx->stream, L.dom.nx,
                     L.dom.ny, G.flags, G.sdf, G.list, G.list_count, G.links);
  FB_HIP(hipGetLastError());
  G.angle = pose.angle_degrees;
  G.valid = true;
  return FB_OK;
}

// Seeds of the postcondition reductions: the minimum at the all-ones key (above every finite key), the
// maxima at key 0 (below the key of every value), the counts at 0.
__global__ void red_seed_kernel(unsigned long long* red) {
  if (threadIdx.x < 6) red[threadIdx.x] = threadIdx.x == 0 ? ~0ull : 0ull;
}

static int reset_red(Lattice& L) {
  // a kernel, not a copy from host memory: queued work never reads a host buffer after this returns
  hipLaunchKernelGGL(red_seed_kernel, dim3(1), dim3(64), 0, L.ctx->stream, L.red);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

static int read_red(Lattice& L, fb_lbm_stats* s) {
  FB_HIP(hipMemcpyAsync(L.red_host, L.red, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, L.ctx->stream));
  FB_HIP(hipStreamSynchronize(L.ctx->stream));
  if (s) {
    s->min_population = dkey_inv(L.red_host[0]);
    s->max_speed = L.red_host[1] ? dkey_inv(L.red_host[1]) : 0.0;
    s->density_excursion = L.red_host[2] ? dkey_inv(L.red_host[2]) : 0.0;
    s->nonfinite_populations = (int64_t)L.red_host[3];
    s->bad_fluid_density = (int64_t)L.red_host[4];
    s->nonfinite_velocity = (int64_t)L.red_host[5];
  }
  return FB_OK;
}

template <class P>
static int advance_impl(Lattice& L, const fb_lbm_params& p, int substeps, const fb_pose* poses, fb_lbm_stats* stats,
                        bool sync_stats, hipEvent_t* ev = nullptr) {
  int rc;
  // Prime f* from the pre-collision state when it is stale or the first pose differs.
  int g_first = L.g;
  bool moved0 = poses && poses[0].angle_degrees != L.G[L.g].angle;
  if (moved0) {
    g_first = other3(L.cvalid ? L.cg : -1, L.g);
    if ((rc = compute_geometry(L, g_first, poses[0]))) return rc;
  }
  int fcur = L.f;
  if (!L.fstar_valid || moved0) {
    int fo = other3(L.cvalid ? L.cf : -1, L.f);
    if ((rc = launch_step<P>(L, p, L.f, true, fo, L.o, L.o, L.g, g_first, false, 0, moved0))) return rc;
    fcur = fo;
    L.f = fo;
    L.g = g_first;
    L.fstar_valid = true;
  }
  if ((rc = reset_red(L))) return rc;
  int pre_out = other3(L.cvalid ? L.cpre : -1, L.pre);
  const bool pairs = pair_ok<typename P::S>(L);
  // the geometry changes after step j (the pose of step j + 1 differs)
  auto moves_after = [&](int j) {
    return poses && j + 1 < substeps && poses[j + 1].angle_degrees != poses[j].angle_degrees;
  };
  for (int k = 0; k < substeps; ++k) {
    bool last = (k == substeps - 1);
    int g_next = L.g;
    bool moving = moves_after(k);
    if (moving) {
      g_next = other3(L.cvalid ? L.cg : -1, L.g);
      if ((rc = compute_geometry(L, g_next, poses[k + 1]))) return rc;
    }
    int fo = other3(L.cvalid ? L.cf : -1, fcur);
    // the outlet history only evolves with a non-periodic x axis (fb/solvers/lbm.py:628-640)
    int oo = L.dom.periodic_x ? L.o : other3(L.cvalid ? L.co : -1, L.o);
    if (ev && k == 0) FB_HIP(hipEventRecord(ev[0], L.ctx->stream));
    if (pairs && !moving && k + 1 < substeps && !moves_after(k + 1)) {
      ++k;
      last = (k == substeps - 1);
      if ((rc = launch_pair<P>(L, p, fcur, fo, L.o, oo, L.g, last, pre_out))) return rc;
    } else if ((rc = launch_step<P>(L, p, fcur, false, fo, L.o, oo, L.g, g_next, last, pre_out, moving))) {
      return rc;
    }
    if (ev && last) FB_HIP(hipEventRecord(ev[1], L.ctx->stream));
    fcur = fo;
    L.f = fo;
    L.o = oo;
    L.g = g_next;
  }
  L.pre = pre_out;
  if (sync_stats) return read_red(L, stats);
  return FB_OK;
}

template <class P>
static int stats_impl(Lattice& L, const fb_lbm_params& p, fb_lbm_stats* stats);

template <class P>
__global__ void __launch_bounds__(256) lbm_stats_kernel(LbmArgs<typename P::S> a) {
  using S = typename P::S;
  using C = typename P::C;
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  double mn = INFINITY, vmax = 0.0, exc = 0.0;
  long long nf = 0, bad = 0, nfu = 0;
  if (x < a.nx) {
    const size_t N = (size_t)a.nx * a.ny;
    const size_t i = (size_t)y * a.nx + x;
    C g[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = ld<S, C>(a.fin + (size_t)d * N + i);
    cell_stats<P>(a, g, !(a.flags[i] & FL_SOLID), mn, vmax, exc, nf, bad, nfu);
  }
  mn = wave_min(mn);
  vmax = wave_max(vmax);
  exc = wave_max(exc);
  nf = wave_sum_ll(nf);
  bad = wave_sum_ll(bad);
  nfu = wave_sum_ll(nfu);
  if ((threadIdx.x & 63) == 0) {
    key_min(&a.red[0], dkey(mn));
    key_max(&a.red[1], dkey(vmax));
    key_max(&a.red[2], dkey(exc));
    if (nf) atomicAdd(&a.red[3], (unsigned long long)nf);
    if (bad) atomicAdd(&a.red[4], (unsigned long long)bad);
    if (nfu) atomicAdd(&a.red[5], (unsigned long long)nfu);
  }
}

template <class P>
static int stats_impl(Lattice& L, const fb_lbm_params& p, fb_lbm_stats* stats) {
  using S = typename P::S;
  int rc;
  if ((rc = reset_red(L))) return rc;
  LbmArgs<S> a{};
  a.nx = L.dom.nx;
  a.ny = L.dom.ny;
  a.fin = (const S*)L.PRE[L.pre];
  a.flags = L.G[L.g].flags;
  a.vscale = p.velocity_scale;
  a.red = L.red;
  hipLaunchKernelGGL(lbm_stats_kernel<P>, dim3((L.dom.nx + 255) / 256, L.dom.ny), dim3(256), 0, L.ctx->stream, a);
  FB_HIP(hipGetLastError());
  return read_red(L, stats);
}

template <class P>
static int rescale_impl(Lattice& L, double ratio) {
  using S = typename P::S;
  int po = other3(L.cvalid ? L.cpre : -1, L.pre);
  int oo = other3(L.cvalid ? L.co : -1, L.o);
  hipLaunchKernelGGL(lbm_rescale_kernel<P>, dim3((L.dom.nx + 255) / 256, L.dom.ny), dim3(256), 0, L.ctx->stream,
                     L.dom.nx, L.dom.ny, (const S*)L.PRE[L.pre], (S*)L.PRE[po], ratio, (S*)L.O[oo]);
  FB_HIP(hipGetLastError());
  L.pre = po;
  L.o = oo;
  L.fstar_valid = false;
  return FB_OK;
}

template <class P>
static int macro_impl(Lattice& L, double vscale, void* rho_host, void* vel_host) {
  using S = typename P::S;
  using B = typename P::B;
  B* d_rho = nullptr;
  B* d_vel = nullptr;
  FB_HIP(dev_alloc((void**)&d_rho, L.N * sizeof(B)));
  FB_HIP(dev_alloc((void**)&d_vel, 2 * L.N * sizeof(B)));
  hipLaunchKernelGGL(lbm_macro_kernel<P>, dim3((L.dom.nx + 255) / 256, L.dom.ny), dim3(256), 0, L.ctx->stream,
                     L.dom.nx, L.dom.ny, (const S*)L.PRE[L.pre], vscale, d_rho, d_vel);
  FB_HIP(hipGetLastError());
  if (rho_host) FB_HIP(hipMemcpyAsync(rho_host, d_rho, L.N * sizeof(B), hipMemcpyDeviceToHost, L.ctx->stream));
  if (vel_host) FB_HIP(hipMemcpyAsync(vel_host, d_vel, 2 * L.N * sizeof(B), hipMemcpyDeviceToHost, L.ctx->stream));
  FB_HIP(hipStreamSynchronize(L.ctx->stream));
  FB_HIP(dev_free(d_rho));
  FB_HIP(dev_free(d_vel));
  return FB_OK;
}

template <class P>
static int set_state_impl(Lattice& L, const void* f_host, const void* outlet_host) {
  using S = typename P::S;
  using B = typename P::B;
  const size_t n9 = L.N * 9;
  B* stage = nullptr;
  FB_HIP(dev_alloc((void**)&stage, n9 * sizeof(B)));
  FB_HIP(hipMemcpyAsync(stage, f_host, n9 * sizeof(B), hipMemcpyHostToDevice, L.ctx->stream));
  hipLaunchKernelGGL(lbm_aos_to_soa<P>, dim3((unsigned)((n9 + 255) / 256)), dim3(256), 0, L.ctx->stream, L.dom.nx,
                     L.dom.ny, stage, (S*)L.PRE[L.pre]);
  FB_HIP(hipGetLastError());
  FB_HIP(hipMemcpyAsync(stage, outlet_host, (size_t)L.dom.ny * 9 * sizeof(B), hipMemcpyHostToDevice, L.ctx->stream));
  hipLaunchKernelGGL(lbm_outlet_convert<P>, dim3((L.dom.ny * 9 + 255) / 256), dim3(256), 0, L.ctx->stream, L.dom.ny,
                     stage, (S*)L.O[L.o], (const S*)nullptr, (B*)nullptr);
  FB_HIP(hipGetLastError());
  FB_HIP(hipStreamSynchronize(L.ctx->stream));
  FB_HIP(dev_free(stage));
  L.fstar_valid = false;
  return FB_OK;
}

template <class P>
static int get_state_impl(Lattice& L, void* f_host, void* outlet_host) {
  using S = typename P::S;
  using B = typename P::B;
  const size_t n9 = L.N * 9;
  B* stage = nullptr;
  FB_HIP(dev_alloc((void**)&stage, n9 * sizeof(B)));
  if (f_host) {
    hipLaunchKernelGGL(lbm_soa_to_aos<P>, dim3((unsigned)((n9 + 255) / 256)), dim3(256), 0, L.ctx->stream, L.dom.nx,
                       L.dom.ny, (const S*)L.PRE[L.pre], stage);
    FB_HIP(hipGetLastError());
    FB_HIP(hipMemcpyAsync(f_host, stage, n9 * sizeof(B), hipMemcpyDeviceToHost, L.ctx->stream));
    FB_HIP(hipStreamSynchronize(L.ctx->stream));
  }
  if (outlet_host) {
    hipLaunchKernelGGL(lbm_outlet_convert<P>, dim3((L.dom.ny * 9 + 255) / 256), dim3(256), 0, L.ctx->stream,
                       L.dom.ny, (const B*)nullptr, (S*)nullptr, (const S*)L.O[L.o], stage);
    FB_HIP(hipGetLastError());
    FB_HIP(hipMemcpyAsync(outlet_host, stage, (size_t)L.dom.ny * 9 * sizeof(B), hipMemcpyDeviceToHost,
                          L.ctx->stream));
    FB_HIP(hipStreamSynchronize(L.ctx->stream));
  }
  FB_HIP(dev_free(stage));
  return FB_OK;
}

template <class P>
static int equilibrium_impl(Lattice& L, const void* rho_host, const void* vel_host) {
  using S = typename P::S;
  using B = typename P::B;
  B* d_rho = nullptr;
  B* d_vel = nullptr;
  FB_HIP(dev_alloc((void**)&d_rho, L.N * sizeof(B)));
  FB_HIP(dev_alloc((void**)&d_vel, 2 * L.N * sizeof(B)));
  FB_HIP(hipMemcpyAsync(d_rho, rho_host, L.N * sizeof(B), hipMemcpyHostToDevice, L.ctx->stream));
  FB_HIP(hipMemcpyAsync(d_vel, vel_host, 2 * L.N * sizeof(B), hipMemcpyHostToDevice, L.ctx->stream));
  int po = other3(L.cvalid ? L.cpre : -1, L.pre);
  int oo = other3(L.cvalid ? L.co : -1, L.o);
  hipLaunchKernelGGL(lbm_equilibrium_kernel<P>, dim3((L.dom.nx + 255) / 256, L.dom.ny), dim3(256), 0, L.ctx->stream,
                     L.dom.nx, L.dom.ny, d_rho, d_vel, (S*)L.PRE[po], (S*)L.O[oo]);
  FB_HIP(hipGetLastError());
  FB_HIP(hipStreamSynchronize(L.ctx->stream));
  FB_HIP(dev_free(d_rho));
  FB_HIP(dev_free(d_vel));
  L.pre = po;
  L.o = oo;
  L.fstar_valid = false;
  return FB_OK;
}

template <class P>
static int uncover_impl(Lattice& L, const fb_pose& pose, const double* feq_u) {
  using S = typename P::S;
  int gi = other3(L.cvalid ? L.cg : -1, L.g);
  int rc = compute_geometry(L, gi, pose);
  if (rc) return rc;
  int po = other3(L.cvalid ? L.cpre : -1, L.pre);
  LbmFeq fq;
  for (int d = 0; d < 9; ++d) fq.v[d] = feq_u[d];
  hipLaunchKernelGGL(lbm_uncover_kernel<P>, dim3((L.dom.nx + 255) / 256, L.dom.ny), dim3(256), 0, L.ctx->stream,
                     L.dom.nx, L.dom.ny, L.G[L.g].flags, L.G[gi].flags, (const S*)L.PRE[L.pre], (S*)L.PRE[po], fq);
  FB_HIP(hipGetLastError());
  L.pre = po;
  L.g = gi;
  L.fstar_valid = false;
  return FB_OK;
}

// Diagnostics and velocity samples of the physical velocity (lattice velocity times vscale) and density,
// formed by the macroscopic kernel into device buffers and reduced there (diag.h).
template <class P>
static int diag_impl(Lattice& L, double vscale, int export_f64, const fb_diag_params* prm, fb_diag* out,
                     int64_t n, const double* points, void* samples) {
  using S = typename P::S;
  using B = typename P::B;
  B* d_rho = nullptr;
  B* d_vel = nullptr;
  FB_HIP(dev_alloc((void**)&d_rho, L.N * sizeof(B)));
  FB_HIP(dev_alloc((void**)&d_vel, 2 * L.N * sizeof(B)));
  hipLaunchKernelGGL(lbm_macro_kernel<P>, dim3((L.dom.nx + 255) / 256, L.dom.ny), dim3(256), 0, L.ctx->stream,
                     L.dom.nx, L.dom.ny, (const S*)L.PRE[L.pre], vscale, d_rho, d_vel);
  int rc = hipGetLastError() == hipSuccess ? FB_OK : FB_ERR_HIP;
  if (rc) set_error("macroscopic kernel launch failed");
  DiagField f;
  f.nx = L.dom.nx;
  f.ny = L.dom.ny;
  f.px = L.dom.periodic_x;
  f.py = L.dom.periodic_y;
  f.x0 = L.dom.x0;
  f.y0 = L.dom.y0;
  f.dx = (L.dom.x1 - L.dom.x0) / L.dom.nx;
  f.dy = (L.dom.y1 - L.dom.y0) / L.dom.ny;
  f.kind = kDiagCells;
  f.storage_f64 = sizeof(B) == sizeof(double);
  f.arith_f64 = export_f64 || f.storage_f64;
  f.u = d_vel;
  f.v = d_vel + 1;
  f.stride = 2;
  f.solid = L.G[L.g].solid;
  if (!rc && prm) {
    f.rho = d_rho;
    f.sdf = L.G[L.g].sdf;
    f.pivot_x = L.foil.pivot_x;
    f.pivot_y = L.foil.pivot_y;
    rc = field_diagnostics(L.ctx, f, *prm, out);
  } else if (!rc) {
    rc = sample_cell_velocity(L.ctx, f, n, points, samples);
  }
  (void)hipStreamSynchronize(L.ctx->stream);
  (void)dev_free(d_rho);
  (void)dev_free(d_vel);
  return rc;
}

#define FB_DISPATCH(prec, fn, ...)                        \
  switch (prec) {                                         \
    case FB_F64: return fn<Prec<FB_F64>>(__VA_ARGS__);     \
    case FB_F32: return fn<Prec<FB_F32>>(__VA_ARGS__);     \
    case FB_F32_MIXED: return fn<Prec<FB_F32_MIXED>>(__VA_ARGS__); \
    case FB_F16_MIXED: return fn<Prec<FB_F16_MIXED>>(__VA_ARGS__); \
    default: set_error("unknown precision %d", prec); return FB_ERR_ARG; \
  }

static int advance_dispatch(Lattice& L, const fb_lbm_params& p, int n, const fb_pose* poses, fb_lbm_stats* s,
                            bool red, hipEvent_t* ev = nullptr) {
  FB_DISPATCH(L.prec, advance_impl, L, p, n, poses, s, red, ev);
}
static int stats_dispatch(Lattice& L, const fb_lbm_params& p, fb_lbm_stats* s) {
  FB_DISPATCH(L.prec, stats_impl, L, p, s);
}
static int rescale_dispatch(Lattice& L, double r) { FB_DISPATCH(L.prec, rescale_impl, L, r); }
static int uncover_dispatch(Lattice& L, const fb_pose& p, const double* f) {
  FB_DISPATCH(L.prec, uncover_impl, L, p, f);
}
static int equilibrium_dispatch(Lattice& L, const void* r, const void* v) {
  FB_DISPATCH(L.prec, equilibrium_impl, L, r, v);
}
static int macro_dispatch(Lattice& L, double v, void* a, void* b) { FB_DISPATCH(L.prec, macro_impl, L, v, a, b); }
static int diag_dispatch(Lattice& L, double v, int e, const fb_diag_params* p, fb_diag* o, int64_t n,
                         const double* pts, void* smp) {
  FB_DISPATCH(L.prec, diag_impl, L, v, e, p, o, n, pts, smp);
}
static int set_state_dispatch(Lattice& L, const void* a, const void* b) { FB_DISPATCH(L.prec, set_state_impl, L, a, b); }
static int get_state_dispatch(Lattice& L, void* a, void* b) { FB_DISPATCH(L.prec, get_state_impl, L, a, b); }

static void free_lattice(Lattice* L) {
  for (int k = 0; k < 3; ++k) {
    if (L->F[k]) (void)dev_free(L->F[k]);
    if (L->R[k]) (void)dev_free(L->R[k]);
    if (L->O[k]) (void)dev_free(L->O[k]);
    if (L->G[k].sdf) (void)dev_free(L->G[k].sdf);
    if (L->G[k].solid) (void)dev_free(L->G[k].solid);
    if (L->G[k].flags) (void)dev_free(L->G[k].flags);
    if (L->G[k].list) (void)dev_free(L->G[k].list);
    if (L->G[k].list_count) (void)dev_free(L->G[k].list_count);
    if (L->G[k].links) (void)dev_free(L->G[k].links);
    if (L->G[k].rows) (void)dev_free(L->G[k].rows);
    if (L->G[k].lidx) (void)dev_free(L->G[k].lidx);
  }
  for (int k = 0; k < 3; ++k)
    if (L->PRE[k]) (void)dev_free(L->PRE[k]);
  if (L->sponge_x) (void)dev_free(L->sponge_x);
  if (L->sponge_y) (void)dev_free(L->sponge_y);
  if (L->red) (void)dev_free(L->red);
  if (L->red_host) (void)host_free(L->red_host);
  if (L->spart) (void)dev_free(L->spart);
  delete L;
}

// ------------------------------------------------------------------ force on the foil

// Momentum exchange on the pre-collision state (build-defined; the reference computes none): for
// every fluid cell x and direction e whose neighbour x + c_e is solid, the link carries
// c_e (f_e(x) + f_opp(e)(x)) into the wall. Lattice units; FP64 sums per block in a fixed order.
template <class P>
__global__ void __launch_bounds__(256) lbm_force_kernel(int nx, int ny, int px, int py,
                                                        const typename P::S* __restrict__ f,
                                                        const uint8_t* __restrict__ flags, double* __restrict__ part) {
  using S = typename P::S;
  using C = typename P::C;
  __shared__ double lds[2][4];
  double fx = 0.0, fy = 0.0;
  const size_t N = (size_t)nx * ny;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t)gridDim.x * 256) {
    if (flags[i] & FL_SOLID) continue;
    const int x = (int)(i % nx), y = (int)(i / nx);
#pragma unroll
    for (int e = 1; e < 9; ++e) {
      int xn = x + D2Q9::cx[e], yn = y + D2Q9::cy[e];
      if (px) xn = (xn + nx) % nx;
      else if (xn < 0 || xn >= nx) continue;
      if (py) yn = (yn + ny) % ny;
      else if (yn < 0 || yn >= ny) continue;
      if (!(flags[(size_t)yn * nx + xn] & FL_SOLID)) continue;
      const int o = D2Q9::opp[e];
      double fe = (double)ld<S, C>(f + (size_t)e * N + i), fo = (double)ld<S, C>(f + (size_t)o * N + i);
      if constexpr (P::shift) {
        fe += D2Q9::w[e];
        fo += D2Q9::w[o];
      }
      fx += D2Q9::cx[e] * (fe + fo);
      fy += D2Q9::cy[e] * (fe + fo);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    fx += __shfl_xor(fx, o, 64);
    fy += __shfl_xor(fy, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    lds[0][threadIdx.x >> 6] = fx;
    lds[1][threadIdx.x >> 6] = fy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ((lds[0][0] + lds[0][1]) + lds[0][2]) + lds[0][3];
    part[2 * blockIdx.x + 1] = ((lds[1][0] + lds[1][1]) + lds[1][2]) + lds[1][3];
  }
}

// Fixed-order sum of n pairs of block partials into out[0..1]: thread t sums pairs t, t + 256, ...
// in order, then a fixed tree combines the 256 thread sums.
__global__ void __launch_bounds__(256) lbm_sum_pairs_kernel(const double* __restrict__ part, int n,
                                                            double* __restrict__ out) {
  __shared__ double lds[2][256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int k = t; k < n; k += 256) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  lds[0][t] = a;
  lds[1][t] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      lds[0][t] += lds[0][t + s];
      lds[1][t] += lds[1][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = lds[0][0];
    out[1] = lds[1][0];
  }
}

template <class P>
static int forces_impl(Lattice& L, double* force) {
  hipStream_t s = L.ctx->stream;
  const unsigned nb = reduction_blocks(L.N);
  double* part = reinterpret_cast<double*>(L.ctx->partials);
  double* out = part + 2 * (size_t)kPartialBlocks;
  hipLaunchKernelGGL(lbm_force_kernel<P>, dim3(nb), dim3(256), 0, s, L.dom.nx, L.dom.ny, L.dom.periodic_x,
                     L.dom.periodic_y, (const typename P::S*)L.PRE[L.pre], (const uint8_t*)L.G[L.g].flags, part);
  hipLaunchKernelGGL(lbm_sum_pairs_kernel, dim3(1), dim3(256), 0, s, part, (int)nb, out);
  FB_HIP(hipGetLastError());
  double h[2];
  FB_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
  FB_HIP(hipStreamSynchronize(s));
  force[0] = h[0];
  force[1] = h[1];
  return FB_OK;
}
static int forces_dispatch(Lattice& L, double* f) { FB_DISPATCH(L.prec, forces_impl, L, f); }

// ------------------------------------------------------------------ batched stepping

struct LbmMember {
  Lattice* L = nullptr;
  fb_lbm_params p{};
  int64_t id = 0;
  int total = 0, remaining = 0;
  Ctx* home = nullptr;
  int list_blocks = 0;  // fixed at join, so the block table and the arguments agree
};

struct LbmBatch {
  Ctx* ctx = nullptr;
  int prec = FB_F32;
  int chunk = 32;
  bool use_graph = true;
  std::vector<LbmMember> m;
  bool changed = true;       // membership changed since the graph was built
  void* d_args = nullptr;    // chunk x members LbmArgs
  size_t args_bytes = 0;
  int* d_start = nullptr;
  size_t start_cap = 0;
  std::vector<unsigned char> h_args, h_prev;
  std::vector<int> h_start;
  int total_blocks = 0;
  hipGraphExec_t exec = nullptr;
};

template <class P>
static int batch_chunk(LbmBatch& B, std::vector<LbmDone>& done) {
  using S = typename P::S;
  const int n = (int)B.m.size();
  if (n == 0) return FB_OK;
  const int K = B.chunk;
  hipStream_t s = B.ctx->stream;
  if (B.changed) {
    B.h_start.assign(n + 1, 0);
    int off = 0;
    for (int i = 0; i < n; ++i) {
      B.h_start[i] = off;
      const Lattice& L = *B.m[i].L;
      const int nb = interior_blocks(L.dom.nx, L.dom.ny, 1) + B.m[i].list_blocks + edge_blocks_for(L.dom.nx, L.dom.ny);
      off += (nb + 7) / 8 * 8;
    }
    B.h_start[n] = off;
    B.total_blocks = off;
    if ((size_t)(n + 1) > B.start_cap) {
      if (B.d_start) FB_HIP(dev_free(B.d_start));
      B.start_cap = (size_t)(n + 1) * 2;
      FB_HIP(dev_alloc((void**)&B.d_start, B.start_cap * sizeof(int)));
    }
    FB_HIP(hipMemcpyAsync(B.d_start, B.h_start.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    const size_t need = (size_t)K * n * sizeof(LbmArgs<S>);
    if (need > B.args_bytes) {
      if (B.d_args) FB_HIP(dev_free(B.d_args));
      B.args_bytes = need * 2;
      FB_HIP(dev_alloc(&B.d_args, B.args_bytes));
    }
    B.h_prev.clear();
    if (B.exec) {
      (void)graph_exec_free(B.exec);
      B.exec = nullptr;
    }
  }
  // arguments of every step of the chunk; a lattice with fewer steps left sits the rest out
  B.h_args.assign((size_t)K * n * sizeof(LbmArgs<S>), 0);
  auto* args = reinterpret_cast<LbmArgs<S>*>(B.h_args.data());
  std::vector<int> pre_out(n);
  for (int i = 0; i < n; ++i) {
    LbmMember& mb = B.m[i];
    Lattice& L = *mb.L;
    int f = L.f, o = L.o;
    pre_out[i] = other3(L.cvalid ? L.cpre : -1, L.pre);
    const int here = std::min(mb.remaining, K);
    for (int j = 0; j < K; ++j) {
      LbmArgs<S>& a = args[(size_t)j * n + i];
      if (j < here) {
        const int fo = other3(L.cvalid ? L.cf : -1, f);
        const int oo = L.dom.periodic_x ? o : other3(L.cvalid ? L.co : -1, o);
        const bool last = j == mb.remaining - 1;
        a = step_args<S>(L, mb.p, f, false, fo, o, oo, L.g, L.g, last, pre_out[i], false);
        f = fo;
        o = oo;
      } else {
        a = step_args<S>(L, mb.p, f, false, f, o, o, L.g, L.g, false, pre_out[i], false);
        a.skip = 1;
      }
      a.list_blocks = mb.list_blocks;
      a.int_rows = 1;
    }
    if (mb.remaining <= K) {
      int rc = reset_red(L);
      if (rc) return rc;
    }
  }
  if (B.h_prev.size() != B.h_args.size() || memcmp(B.h_prev.data(), B.h_args.data(), B.h_args.size()) != 0) {
    FB_HIP(hipMemcpyAsync(B.d_args, B.h_args.data(), B.h_args.size(), hipMemcpyHostToDevice, s));
    // the copy reads pageable memory: the host buffer may be reused once the call returns
    B.h_prev = B.h_args;
  }
  auto launch_all = [&](Launcher& L) {
    for (int j = 0; j < K; ++j)
      L(lbm_batch_kernel<P>, dim3(B.total_blocks), dim3(256), 0, (const LbmArgs<S>*)B.d_args + (size_t)j * n,
        (const int*)B.d_start, n);
  };
  if (B.use_graph) {
    if (!B.exec) {
      // the chunk's graph is built node by node (no stream capture)
      hipGraph_t graph = nullptr;
      FB_HIP(hipGraphCreate(&graph, 0));
      Launcher L{s, graph};
      launch_all(L);
      hipError_t e = L.status();
      if (e == hipSuccess) {
        std::lock_guard<std::recursive_mutex> guard(hip_lock());
        e = hipGraphInstantiate(&B.exec, graph, nullptr, nullptr, 0);
      }
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        B.exec = nullptr;
        set_error("lattice batch graph failed: %s", hipGetErrorString(e));
        return FB_ERR_HIP;
      }
    }
    FB_HIP(hipGraphLaunch(B.exec, s));
  } else {
    Launcher L{s};
    launch_all(L);
    FB_HIP(L.status());
  }
  B.changed = false;
  FB_HIP(hipStreamSynchronize(s));
  // bookkeeping: buffer indices after the chunk, finished lattices leave
  std::vector<LbmMember> keep;
  keep.reserve(n);
  for (int i = 0; i < n; ++i) {
    LbmMember& mb = B.m[i];
    Lattice& L = *mb.L;
    const int here = std::min(mb.remaining, K);
    for (int j = 0; j < here; ++j) {
      L.f = other3(L.cvalid ? L.cf : -1, L.f);
      if (!L.dom.periodic_x) L.o = other3(L.cvalid ? L.co : -1, L.o);
    }
    mb.remaining -= here;
    if (mb.remaining == 0) {
      L.pre = pre_out[i];
      LbmDone d;
      d.id = mb.id;
      d.steps = mb.total;
      d.status = read_red(L, &d.stats);
      L.ctx = mb.home;
      done.push_back(d);
      B.changed = true;
    } else {
      keep.push_back(mb);
    }
  }
  B.m.swap(keep);
  return FB_OK;
}

static int batch_chunk_dispatch(LbmBatch& B, std::vector<LbmDone>& done) { FB_DISPATCH(B.prec, batch_chunk, B, done); }

template <class P>
static int batch_prime(Lattice& L, const fb_lbm_params& p) {
  // f* from the pre-collision state, as the first step of an advance does
  if (L.fstar_valid) return FB_OK;
  const int fo = other3(L.cvalid ? L.cf : -1, L.f);
  int rc = launch_step<P>(L, p, L.f, true, fo, L.o, L.o, L.g, L.g, false, 0, false);
  if (rc) return rc;
  L.f = fo;
  L.fstar_valid = true;
  return FB_OK;
}
static int batch_prime_dispatch(Lattice& L, const fb_lbm_params& p) { FB_DISPATCH(L.prec, batch_prime, L, p); }

int lbm_batch_create(Ctx* c, int prec, int chunk, int use_graph, LbmBatch** out) {
  FB_REQUIRE(c && out, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(prec >= FB_F64 && prec <= FB_F16_MIXED, FB_ERR_ARG, "unknown precision %d", prec);
  FB_REQUIRE(chunk >= 2 && chunk % 2 == 0 && chunk <= 1024, FB_ERR_ARG, "chunk must be even and in [2, 1024]");
  auto* b = new LbmBatch();
  b->ctx = c;
  b->prec = prec;
  b->chunk = chunk;
  b->use_graph = use_graph != 0;
  *out = b;
  return FB_OK;
}

void lbm_batch_destroy(LbmBatch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->ctx->stream);
  for (auto& mb : b->m) mb.L->ctx = mb.home;
  if (b->exec) (void)graph_exec_free(b->exec);
  if (b->d_args) (void)dev_free(b->d_args);
  if (b->d_start) (void)dev_free(b->d_start);
  delete b;
}

int lbm_batch_add(LbmBatch* b, fb_lbm* lattice, const fb_lbm_params& p, int steps, int64_t id) {
  FB_REQUIRE(b && lattice, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(steps >= 1, FB_ERR_ARG, "steps must be positive");
  auto* L = reinterpret_cast<Lattice*>(lattice);
  FB_REQUIRE(L->prec == b->prec, FB_ERR_ARG, "lattice precision %d differs from the batch's %d", L->prec, b->prec);
  FB_REQUIRE(L->G[L->g].valid, FB_ERR_STATE, "no geometry has been set");
  for (auto& mb : b->m) FB_REQUIRE(mb.L != L, FB_ERR_STATE, "lattice is already in the batch");
  LbmMember mb;
  mb.L = L;
  mb.p = p;
  mb.id = id;
  mb.total = steps;
  mb.remaining = steps;
  mb.home = L->ctx;
  mb.list_blocks = list_blocks(*L);
  L->ctx = b->ctx;
  int rc = batch_prime_dispatch(*L, p);
  if (rc) {
    L->ctx = mb.home;
    return rc;
  }
  b->m.push_back(mb);
  b->changed = true;
  return FB_OK;
}

int lbm_batch_members(const LbmBatch* b) { return b ? (int)b->m.size() : 0; }
int lbm_handle_precision(const fb_lbm* lattice) { return reinterpret_cast<const Lattice*>(lattice)->prec; }
hipError_t lbm_handle_sync(fb_lbm* lattice) {
  return hipStreamSynchronize(reinterpret_cast<Lattice*>(lattice)->ctx->stream);
}
int lbm_batch_prec(const LbmBatch* b) { return b ? b->prec : -1; }
int64_t lbm_batch_cells(const LbmBatch* b) {
  int64_t c = 0;
  if (b)
    for (auto& mb : b->m) c += (int64_t)mb.L->N;
  return c;
}
int lbm_batch_run_chunk(LbmBatch* b, std::vector<LbmDone>& done) {
  FB_REQUIRE(b, FB_ERR_ARG, "NULL batch");
  FB_HIP(hipSetDevice(b->ctx->device));
  return batch_chunk_dispatch(*b, done);
}

}  // namespace fb

using fb::Lattice;

extern "C" {

int fb_lbm_create(fb_ctx* ctx, const fb_domain* dom, const fb_foil* foil, int precision, int channel_walls,
                  const double* sponge_y, const double* sponge_x, fb_lbm** out) {
  FB_REQUIRE(ctx && dom && foil && sponge_y && sponge_x && out, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(dom->nx >= 4 && dom->ny >= 4, FB_ERR_ARG, "each resolution axis must contain at least four cells");
  FB_REQUIRE(precision >= FB_F64 && precision <= FB_F16_MIXED, FB_ERR_ARG, "unknown precision %d", precision);
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  auto* L = new Lattice();
  L->ctx = c;
  L->dom = *dom;
  L->foil = *foil;
  L->prec = precision;
  L->channel = channel_walls;
  L->N = (size_t)dom->nx * dom->ny;
  L->ssize = precision == FB_F64 ? 8 : (precision == FB_F16_MIXED ? 2 : 4);
  L->bsize = precision == FB_F64 ? 8 : 4;
  const size_t N = L->N;
  auto fail = [&](hipError_t e) {
    fb::set_error("device allocation failed: %s", hipGetErrorString(e));
    fb::free_lattice(L);
    return FB_ERR_NOMEM;
  };
  hipError_t e;
  for (int k = 0; k < 3; ++k) {
    if ((e = dev_alloc(&L->F[k], 9 * N * L->ssize))) return fail(e);
    if ((e = dev_alloc(&L->R[k], N * L->ssize))) return fail(e);
    if ((e = dev_alloc(&L->O[k], (size_t)dom->ny * 9 * L->ssize))) return fail(e);
    if ((e = dev_alloc((void**)&L->G[k].sdf, N * sizeof(double)))) return fail(e);
    if ((e = dev_alloc((void**)&L->G[k].solid, N))) return fail(e);
    if ((e = dev_alloc((void**)&L->G[k].flags, N))) return fail(e);
    if ((e = dev_alloc((void**)&L->G[k].list, N * sizeof(int)))) return fail(e);
    if ((e = dev_alloc((void**)&L->G[k].list_count, sizeof(int)))) return fail(e);
    if ((e = dev_alloc((void**)&L->G[k].links, N * 8 * sizeof(double)))) return fail(e);
    if ((e = dev_alloc((void**)&L->G[k].rows, 2 * (size_t)dom->ny * sizeof(int)))) return fail(e);
    if ((e = dev_alloc((void**)&L->G[k].lidx, N * sizeof(int)))) return fail(e);
    if ((e = hipMemsetAsync(L->R[k], 0, N * L->ssize, c->stream))) return fail(e);
  }
  for (int k = 0; k < 3; ++k)
    if ((e = dev_alloc(&L->PRE[k], 9 * N * L->ssize))) return fail(e);
  if ((e = dev_alloc((void**)&L->sponge_y, dom->ny * sizeof(double)))) return fail(e);
  if ((e = dev_alloc((void**)&L->sponge_x, dom->nx * sizeof(double)))) return fail(e);
  if ((e = dev_alloc((void**)&L->red, 8 * sizeof(unsigned long long)))) return fail(e);
  {
    // statistics slots for the largest single-step grid: one interior row per block, the most list blocks
    L->spart_blocks = fb::edge_blocks_for(dom->nx, dom->ny) + 1024 + fb::interior_blocks(dom->nx, dom->ny, 1);
    if ((e = dev_alloc((void**)&L->spart, 6 * (size_t)L->spart_blocks * sizeof(unsigned long long)))) return fail(e);
  }
  if ((e = host_alloc((void**)&L->red_host, 8 * sizeof(unsigned long long)))) return fail(e);
  // stream-ordered copies: a legacy-stream copy would wait on every other thread's stream
  if ((e = hipMemcpyAsync(L->sponge_y, sponge_y, dom->ny * sizeof(double), hipMemcpyHostToDevice, c->stream)))
    return fail(e);
  if ((e = hipMemcpyAsync(L->sponge_x, sponge_x, dom->nx * sizeof(double), hipMemcpyHostToDevice, c->stream)))
    return fail(e);
  if ((e = hipStreamSynchronize(c->stream))) return fail(e);
  *out = reinterpret_cast<fb_lbm*>(L);
  return FB_OK;
}

int fb_lbm_destroy(fb_lbm* lbm) {
  if (!lbm) return FB_OK;
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_HIP(hipSetDevice(L->ctx->device));
  FB_HIP(hipStreamSynchronize(L->ctx->stream));
  fb::free_lattice(L);
  return FB_OK;
}

int fb_lbm_set_state(fb_lbm* lbm, const void* f_host, const void* outlet_host) {
  FB_REQUIRE(lbm && f_host && outlet_host, FB_ERR_ARG, "NULL argument");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::set_state_dispatch(*L, f_host, outlet_host);
}

int fb_lbm_get_state(fb_lbm* lbm, void* f_host, void* outlet_host) {
  FB_REQUIRE(lbm, FB_ERR_ARG, "NULL handle");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::get_state_dispatch(*L, f_host, outlet_host);
}

int fb_lbm_set_pose(fb_lbm* lbm, const fb_pose* pose) {
  FB_REQUIRE(lbm && pose, FB_ERR_ARG, "NULL argument");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_HIP(hipSetDevice(L->ctx->device));
  int gi = fb::other3(L->cvalid ? L->cg : -1, L->g);
  int rc = fb::compute_geometry(*L, gi, *pose);
  if (rc) return rc;
  L->g = gi;
  L->fstar_valid = false;
  return FB_OK;
}

int fb_lbm_update_pose(fb_lbm* lbm, const fb_pose* pose, const fb_lbm_params* params) {
  FB_REQUIRE(lbm && pose && params, FB_ERR_ARG, "NULL argument");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_REQUIRE(L->G[L->g].valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(L->ctx->device));
  if (pose->angle_degrees == L->G[L->g].angle) return FB_OK;
  return fb::uncover_dispatch(*L, *pose, params->feq_uncovered);
}

int fb_lbm_get_geometry(fb_lbm* lbm, double* sdf_host, uint8_t* solid_host) {
  FB_REQUIRE(lbm, FB_ERR_ARG, "NULL handle");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_REQUIRE(L->G[L->g].valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(L->ctx->device));
  if (sdf_host)
    FB_HIP(hipMemcpyAsync(sdf_host, L->G[L->g].sdf, L->N * sizeof(double), hipMemcpyDeviceToHost, L->ctx->stream));
  if (solid_host)
    FB_HIP(hipMemcpyAsync(solid_host, L->G[L->g].solid, L->N, hipMemcpyDeviceToHost, L->ctx->stream));
  FB_HIP(hipStreamSynchronize(L->ctx->stream));
  return FB_OK;
}

int fb_lbm_checkpoint(fb_lbm* lbm) {
  FB_REQUIRE(lbm, FB_ERR_ARG, "NULL handle");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  L->cf = L->f;
  L->co = L->o;
  L->cpre = L->pre;
  L->cg = L->g;
  L->cfstar = L->fstar_valid;
  L->cvalid = true;
  return FB_OK;
}

int fb_lbm_rollback(fb_lbm* lbm) {
  FB_REQUIRE(lbm, FB_ERR_ARG, "NULL handle");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_REQUIRE(L->cvalid, FB_ERR_STATE, "no checkpoint to roll back to");
  FB_HIP(hipSetDevice(L->ctx->device));
  FB_HIP(hipStreamSynchronize(L->ctx->stream));
  L->f = L->cf;
  L->o = L->co;
  L->pre = L->cpre;
  L->g = L->cg;
  L->fstar_valid = L->cfstar;
  return FB_OK;
}

int fb_lbm_rescale(fb_lbm* lbm, double ratio) {
  FB_REQUIRE(lbm, FB_ERR_ARG, "NULL handle");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::rescale_dispatch(*L, ratio);
}

int fb_lbm_set_equilibrium(fb_lbm* lbm, const void* rho_host, const void* vel_host) {
  FB_REQUIRE(lbm && rho_host && vel_host, FB_ERR_ARG, "NULL argument");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::equilibrium_dispatch(*L, rho_host, vel_host);
}

int fb_lbm_advance(fb_lbm* lbm, const fb_lbm_params* params, int substeps, const fb_pose* poses,
                   fb_lbm_stats* stats) {
  FB_REQUIRE(lbm && params && poses, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(substeps >= 1, FB_ERR_ARG, "substeps must be positive");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_REQUIRE(L->G[L->g].valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::advance_dispatch(*L, *params, substeps, poses, stats, true);
}

int fb_lbm_stats_now(fb_lbm* lbm, const fb_lbm_params* params, fb_lbm_stats* stats) {
  FB_REQUIRE(lbm && params && stats, FB_ERR_ARG, "NULL argument");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::stats_dispatch(*L, *params, stats);
}

int fb_lbm_macroscopic(fb_lbm* lbm, double velocity_scale, void* rho_host, void* vel_host) {
  FB_REQUIRE(lbm, FB_ERR_ARG, "NULL handle");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::macro_dispatch(*L, velocity_scale, rho_host, vel_host);
}

int fb_lbm_time_fixed(fb_lbm* lbm, const fb_lbm_params* params, int steps, double* elapsed_ms) {
  FB_REQUIRE(lbm && params && elapsed_ms, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(steps >= 1 && steps <= 100000000, FB_ERR_ARG, "steps out of range");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_REQUIRE(L->G[L->g].valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(L->ctx->device));
  // Prime outside the timed region so every timed launch is a lattice step.
  if (!L->fstar_valid) {
    int rc = fb::advance_dispatch(*L, *params, 1, nullptr, nullptr, false);
    if (rc) return rc;
  }
  hipEvent_t ev[2];
  FB_HIP(hipEventCreate(&ev[0]));
  FB_HIP(hipEventCreate(&ev[1]));
  FB_HIP(hipStreamSynchronize(L->ctx->stream));
  // One event pair around the back-to-back launches: events between launches would add gaps.
  int rc = fb::advance_dispatch(*L, *params, steps, nullptr, nullptr, false, ev);
  if (rc == FB_OK) {
    FB_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    FB_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    *elapsed_ms = ms;
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return rc;
}

int fb_lbm_run_batch(fb_lbm* const* lattices, const fb_lbm_params* params, int n, int steps, int chunk,
                     fb_lbm_stats* stats) {
  FB_REQUIRE(lattices && params && n >= 1, FB_ERR_ARG, "NULL argument or empty batch");
  FB_REQUIRE(steps >= 1, FB_ERR_ARG, "steps must be positive");
  for (int k = 0; k < n; ++k) FB_REQUIRE(lattices[k] != nullptr, FB_ERR_ARG, "lattice %d is NULL", k);
  auto* L0 = reinterpret_cast<Lattice*>(lattices[0]);
  FB_HIP(hipSetDevice(L0->ctx->device));
  for (int k = 0; k < n; ++k) FB_HIP(fb::lbm_handle_sync(lattices[k]));
  fb::LbmBatch* b = nullptr;
  int rc = fb::lbm_batch_create(L0->ctx, L0->prec, chunk, 1, &b);
  if (rc) return rc;
  for (int k = 0; k < n && rc == FB_OK; ++k) rc = fb::lbm_batch_add(b, lattices[k], params[k], steps, k);
  std::vector<fb::LbmDone> done;
  while (rc == FB_OK && fb::lbm_batch_members(b) > 0) rc = fb::lbm_batch_run_chunk(b, done);
  fb::lbm_batch_destroy(b);
  if (rc) return rc;
  for (auto& d : done) {
    if (d.status) return d.status;
    if (stats) stats[d.id] = d.stats;
  }
  return FB_OK;
}

int fb_lbm_forces(fb_lbm* lbm, double* force) {
  FB_REQUIRE(lbm && force, FB_ERR_ARG, "NULL argument");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_REQUIRE(L->G[L->g].valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::forces_dispatch(*L, force);
}

int fb_lbm_run_fixed(fb_lbm* lbm, const fb_lbm_params* params, int steps) {
  FB_REQUIRE(lbm && params, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(steps >= 1, FB_ERR_ARG, "steps must be positive");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_REQUIRE(L->G[L->g].valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::advance_dispatch(*L, *params, steps, nullptr, nullptr, false);
}

int fb_lbm_diagnostics(fb_lbm* lbm, double velocity_scale, int export_f64, const fb_diag_params* params,
                       fb_diag* out) {
  FB_REQUIRE(lbm && params && out, FB_ERR_ARG, "NULL argument");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_REQUIRE(L->G[L->g].valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(L->ctx->device));
  *out = fb_diag{};
  return fb::diag_dispatch(*L, velocity_scale, export_f64, params, out, 0, nullptr, nullptr);
}

int fb_lbm_sample_velocity(fb_lbm* lbm, double velocity_scale, int export_f64, int64_t n, const double* points,
                           void* out) {
  FB_REQUIRE(lbm, FB_ERR_ARG, "NULL handle");
  auto* L = reinterpret_cast<Lattice*>(lbm);
  FB_REQUIRE(L->G[L->g].valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(L->ctx->device));
  return fb::diag_dispatch(*L, velocity_scale, export_f64, nullptr, nullptr, n, points, out);
}

}  // extern "C"
