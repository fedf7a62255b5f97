// Blended PIC/FLIP on gfx950: the particle half of fb/solvers/pic_flip.py (seeding, collisions,
// swept collisions, midpoint advection, periodic wrap, escape respawn) and the quadratic
// B-spline transfers of kern/pic.py:105-205 behind fb/solvers/_numba_adapter.py:108-160.
// The grid half (implicit viscosity on faces, projection) is the MAC grid of mac.hip.
//
// Layout: positions and velocities (P, 2) interleaved in the storage dtype, like the
// reference arrays; the grid velocity as two (ny, nx) cell planes.
//
// P2G is a per-face gather over particles sorted by cell (counting sort, then each cell's
// list ordered by particle index), so every face sums its contributions in a fixed order
// and no floating-point atomics are used. Each addition follows the numba kernel: the FP64
// contribution is added to the dtype accumulator in FP64 and rounded back (kern/pic.py:144-145).
// G2P is a per-particle gather in the reference's fixed 3 x 3 order with FP64 accumulation.
#include <algorithm>
#include <cmath>
#include <vector>

#include "diag.h"
#include "mac.h"

namespace fb {

#pragma clang fp contract(off)

template <class T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

template <class T>
__device__ __forceinline__ void ld2(const T* __restrict__ a, size_t i, T& x, T& y) {
  auto v = reinterpret_cast<const typename Vec2<T>::type*>(a)[i];
  x = v.x;
  y = v.y;
}
template <class T>
__device__ __forceinline__ void st2(T* __restrict__ a, size_t i, T x, T y) {
  typename Vec2<T>::type v;
  v.x = x;
  v.y = y;
  reinterpret_cast<typename Vec2<T>::type*>(a)[i] = v;
}

struct PicGeom {
  int nx, ny, px, py;
  double x0, y0, x1, y1, dx, dy;
};

inline PicGeom make_pic_geom(const fb_domain& d) {
  PicGeom g;
  g.nx = d.nx;
  g.ny = d.ny;
  g.px = d.periodic_x;
  g.py = d.periodic_y;
  g.x0 = d.x0;
  g.y0 = d.y0;
  g.x1 = d.x1;
  g.y1 = d.y1;
  g.dx = (d.x1 - d.x0) / d.nx;
  g.dy = (d.y1 - d.y0) / d.ny;
  return g;
}

// One face component of the transfers (fb/solvers/_numba_adapter.py:126-133, 150-157).
struct Comp {
  int W, H;    // stored width and height
  int uw, uh;  // unique width and height (a periodic axis drops its duplicate endpoint)
  int dupx, dupy;
  double gxo, gyo;
};

__host__ __device__ inline Comp comp_u(const PicGeom& g) {
  Comp c;
  c.W = g.nx + 1;
  c.H = g.ny;
  c.dupx = g.px;
  c.dupy = 0;
  c.uw = c.W - c.dupx;
  c.uh = c.H;
  c.gxo = 0.0;
  c.gyo = -0.5;
  return c;
}

__host__ __device__ inline Comp comp_v(const PicGeom& g) {
  Comp c;
  c.W = g.nx;
  c.H = g.ny + 1;
  c.dupx = 0;
  c.dupy = g.py;
  c.uw = c.W;
  c.uh = c.H - c.dupy;
  c.gxo = -0.5;
  c.gyo = 0.0;
  return c;
}

// Quadratic B-spline (kern/pic.py:11-19).
__device__ __forceinline__ double bspline(double d) {
  double a = fabs(d);
  if (a < 0.5) return 0.75 - a * a;
  if (a < 1.5) {
    double e = 1.5 - a;
    return 0.5 * e * e;
  }
  return 0.0;
}

// Python's integer modulo.
__device__ __forceinline__ int pmod(int a, int n) {
  int m = a % n;
  return m < 0 ? m + n : m;
}

// int(np.floor(v)); values beyond +-1e9 are clamped (they lie outside every face range).
__device__ __forceinline__ int ifloor(double v) {
  double f = floor(v);
  f = fmin(fmax(f, -1.0e9), 1.0e9);
  return (int)f;
}

// Grid coordinates of a point before the component's face offset: ((x - x0) / dx, (y - y0) / dy),
// the part of kern/pic.py:187-188 that the u and v gathers at one point share (the same FP64
// quotients the reference computes once per component).
__device__ __forceinline__ void grid_coords(const PicGeom& g, double x, double y, double& tx, double& ty) {
  tx = (x - g.x0) / g.dx;
  ty = (y - g.y0) / g.dy;
}

// Gather of one component at the point of grid coordinates (tx, ty) (kern/pic.py:164-205): FP64
// weights and sums in the fixed 3 x 3 order. The three weights per axis are evaluated once (the
// reference evaluates the x weights again for every row; the values are the same). With `q` and
// `p` given, also gathers the dtype difference q - p with the same weights (the FLIP increment of
// fb/solvers/pic_flip.py:822-826).
template <class T, bool DIFF>
__device__ __forceinline__ void gather(const PicGeom& g, const Comp& c, const T* __restrict__ q,
                                       const T* __restrict__ p, double tx, double ty, double& out_q,
                                       double& out_d) {
  const double gx = tx + c.gxo;
  const double gy = ty + c.gyo;
  const int bx = ifloor(gx - 0.5), by = ifloor(gy - 0.5);
  double wxs[3], wys[3];
#pragma unroll
  for (int o = 0; o < 3; ++o) {
    wxs[o] = bspline(gx - (double)(bx + o));
    wys[o] = bspline(gy - (double)(by + o));
  }
  double vq = 0.0, vd = 0.0, ws = 0.0;
#pragma unroll
  for (int oy = 0; oy < 3; ++oy) {
    const int sy = by + oy;
    const int ty_ = g.py ? pmod(sy, c.uh) : sy;
    if (ty_ < 0 || ty_ >= c.H) continue;
#pragma unroll
    for (int ox = 0; ox < 3; ++ox) {
      const int sx = bx + ox;
      const int tx_ = g.px ? pmod(sx, c.uw) : sx;
      if (tx_ < 0 || tx_ >= c.W) continue;
      const double w = wys[oy] * wxs[ox];
      const size_t k = (size_t)ty_ * c.W + tx_;
      const T qv = q[k];
      vq += w * (double)qv;
      if constexpr (DIFF) vd += w * (double)(T)(qv - p[k]);
      ws += w;
    }
  }
  const double den = fmax(ws, 1.0e-12);
  out_q = vq / den;
  if constexpr (DIFF) out_d = vd / den;
}

// Home cell used to bin particles for the P2G gather: the cell containing the particle,
// wrapped on periodic axes and clamped otherwise.
__device__ __forceinline__ int bin_cell(const PicGeom& g, double x, double y) {
  int cx = ifloor((x - g.x0) / g.dx);
  int cy = ifloor((y - g.y0) / g.dy);
  cx = g.px ? pmod(cx, g.nx) : min(max(cx, 0), g.nx - 1);
  cy = g.py ? pmod(cy, g.ny) : min(max(cy, 0), g.ny - 1);
  return cy * g.nx + cx;
}

// Cell-centred transfers (kern/pic.py:22-102).
// Stencil base of a coordinate: grid coordinate g = (x - x0) / d - 0.5, base floor(g - 0.5).
__device__ __forceinline__ double cell_coord(double x, double x0, double d) { return (x - x0) / d - 0.5; }

// Target of stencil index s along an axis of n cells.
__device__ __forceinline__ int cell_target(int s, int n, int periodic) {
  return periodic ? pmod(s, n) : min(max(s, 0), n - 1);
}

// Bin of a stencil base along an axis: wrapped on a periodic axis (n bins); otherwise the base
// clamped to [-1, n - 1] and shifted by one (n + 1 bins). A base below -1 reaches only cell 0 and
// one above n - 2 only cell n - 1, so each end bin holds every particle beyond it and the cells
// that can receive from a bin are those of its bases.
__device__ __forceinline__ int base_bin(int b, int n, int periodic) {
  return periodic ? pmod(b, n) : min(max(b, -1), n - 1) + 1;
}
__host__ __device__ inline int base_bins(int n, int periodic) { return periodic ? n : n + 1; }

__device__ __forceinline__ long long wave_sum_i(long long v) { return wave_sum_ll(v); }

// ------------------------------------------------------------------ PCG32 (fb/core/rng.py)

struct Pcg {
  unsigned long long state, inc;
};

constexpr unsigned long long kPcgMult = 6364136223846793005ull;

__host__ __device__ inline unsigned pcg_output(unsigned long long old) {
  unsigned xs = (unsigned)(((old >> 18) ^ old) >> 27);
  unsigned rot = (unsigned)(old >> 59);
  return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

// State after `delta` steps (the logarithmic jump-ahead of O'Neill's PCG paper, section 4.3.1).
__host__ __device__ inline unsigned long long pcg_advance(unsigned long long state, unsigned long long inc,
                                                          unsigned long long delta) {
  unsigned long long cur_mult = kPcgMult, cur_plus = inc, acc_mult = 1ull, acc_plus = 0ull;
  while (delta > 0) {
    if (delta & 1ull) {
      acc_mult *= cur_mult;
      acc_plus = acc_plus * cur_mult + cur_plus;
    }
    cur_plus = (cur_mult + 1ull) * cur_plus;
    cur_mult *= cur_mult;
    delta >>= 1;
  }
  return acc_mult * state + acc_plus;
}

// float32(u32) * float32(2^-32) (fb/core/rng.py:46-52).
__device__ __forceinline__ float pcg_unit(unsigned u) { return (float)u * (1.0f / 4294967296.0f); }

// Draw number `k` (0-based) after the state `s`.
__device__ __forceinline__ float pcg_draw(const Pcg& s, unsigned long long k) {
  return pcg_unit(pcg_output(pcg_advance(s.state, s.inc, k)));
}

// uniform(low, high) of fb/core/rng.py:54-55: float32 arithmetic with Python floats rounded.
__device__ __forceinline__ float pcg_uniform(float r, double low, double high) {
  return (float)low + (float)(high - low) * r;
}

// ------------------------------------------------------------------ scan

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Exclusive scan of 1024 ints per block; block totals into bsum.
__global__ void __launch_bounds__(256) scan_block_kernel(const int* __restrict__ in, int* __restrict__ out, int n,
                                                         int* __restrict__ bsum) {
  __shared__ int wsum[4];
  const int base = blockIdx.x * 1024 + threadIdx.x * 4;
  int v[4];
  int s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int incl = wave_incl_scan(s);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int off = 0;
  for (int j = 0; j < w; ++j) off += wsum[j];
  int run = off + incl - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = off + incl;
}

// Exclusive scan of the block totals in place (one block); the grand total into *total.
__global__ void __launch_bounds__(256) scan_sums_kernel(int* __restrict__ bsum, int nb, int* __restrict__ total) {
  __shared__ int wsum[4];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < nb; c0 += 1024) {
    const int base = c0 + threadIdx.x * 4;
    int v[4];
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = (base + k < nb) ? bsum[base + k] : 0;
      s += v[k];
    }
    int incl = wave_incl_scan(s);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int off = carry;
    for (int j = 0; j < w; ++j) off += wsum[j];
    int run = off + incl - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (base + k < nb) bsum[base + k] = run;
      run += v[k];
    }
    __syncthreads();
    if (threadIdx.x == 255) carry = off + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256) scan_add_kernel(int* __restrict__ out, int n, const int* __restrict__ bsum,
                                                       const int* __restrict__ total) {
  const int base = blockIdx.x * 1024;
  const int add = bsum[blockIdx.x];
  for (int k = threadIdx.x; k < 1024; k += 256)
    if (base + k < n) out[base + k] += add;
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = *total;
}

__global__ void __launch_bounds__(256) fill_int_kernel(int* __restrict__ a, int n, int v) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = v;
}

// ------------------------------------------------------------------ sort by cell

// BASE: bin by cell-centred stencil base (the bins of p2c_kernel) instead of the home cell.
template <class T, bool BASE>
__global__ void __launch_bounds__(256) bin_count_kernel(PicGeom g, int n, const T* __restrict__ pos,
                                                        int* __restrict__ cellof, int* __restrict__ count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool act = i < n;
  int c = 0;
  if (act) {
    T x, y;
    ld2(pos, i, x, y);
    if constexpr (BASE) {
      const int bx = ifloor(cell_coord((double)x, g.x0, g.dx) - 0.5);
      const int by = ifloor(cell_coord((double)y, g.y0, g.dy) - 0.5);
      c = base_bin(by, g.ny, g.py) * base_bins(g.nx, g.px) + base_bin(bx, g.nx, g.px);
    } else {
      c = bin_cell(g, (double)x, (double)y);
    }
    cellof[i] = c;
  }
  wave_run_add(count, c, act);
}

__global__ void __launch_bounds__(256) bin_scatter_kernel(int n, const int* __restrict__ cellof,
                                                          const int* __restrict__ start, int* __restrict__ cursor,
                                                          int* __restrict__ sidx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool act = i < n;
  const int c = act ? cellof[i] : 0;
  const int slot = wave_run_claim(cursor, c, act);
  if (act) sidx[start[c] + slot] = i;
}

// Order each cell's particles by index (insertion sort; cells hold a handful of particles).
__global__ void __launch_bounds__(256) cell_sort_kernel(int ncell, const int* __restrict__ start,
                                                        int* __restrict__ sidx) {
  int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncell) return;
  int a = start[c], b = start[c + 1];
  for (int i = a + 1; i < b; ++i) {
    int key = sidx[i];
    int j = i - 1;
    while (j >= a && sidx[j] > key) {
      sidx[j + 1] = sidx[j];
      --j;
    }
    sidx[j + 1] = key;
  }
}

template <class T>
__global__ void __launch_bounds__(256) sorted_copy_kernel(int n, const int* __restrict__ sidx,
                                                          const T* __restrict__ pos, const T* __restrict__ vel,
                                                          T* __restrict__ spos, T* __restrict__ sval) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int p = sidx[k];
  T x, y;
  ld2(pos, p, x, y);
  st2(spos, k, x, y);
  ld2(vel, p, x, y);
  st2(sval, k, x, y);
}

// ------------------------------------------------------------------ P2G

template <class T>
struct P2GArgs {
  PicGeom g;
  Comp c;
  int comp;  // 0: u faces from velocity x, 1: v faces from velocity y
  const int* start;
  const T* spos;
  const T* sval;
  const T* fallback;
  T* out;
  unsigned long long* unsupported;
};

// scatter_face_component_kernel (kern/pic.py:105-161) as a gather: face (fx, fy) visits the
// 4 x 4 cells whose particles can reach it and keeps those whose 3 x 3 support contains it.
template <class T>
__global__ void __launch_bounds__(256) p2g_kernel(P2GArgs<T> a) {
  const Comp& c = a.c;
  const PicGeom& g = a.g;
  const int tx = blockIdx.x * 256 + threadIdx.x;
  const int ty = blockIdx.y;
  long long uns = 0;
  if (tx < c.W && ty < c.H) {
    const bool dupc = c.dupx && tx == c.W - 1;
    const bool dupr = c.dupy && ty == c.H - 1;
    const int fx = dupc ? 0 : tx;
    const int fy = dupr ? 0 : ty;
    T wsum = (T)0, msum = (T)0;
    for (int wy = -2; wy <= 1; ++wy) {
      int cy = fy + wy;
      if (g.py) cy = pmod(cy, g.ny);
      else if (cy < 0 || cy >= g.ny) continue;
      for (int wx = -2; wx <= 1; ++wx) {
        int cx = fx + wx;
        if (g.px) cx = pmod(cx, g.nx);
        else if (cx < 0 || cx >= g.nx) continue;
        const int cell = cy * g.nx + cx;
        const int k1 = a.start[cell + 1];
        for (int k = a.start[cell]; k < k1; ++k) {
          T x, y;
          ld2(a.spos, k, x, y);
          double gx = ((double)x - g.x0) / g.dx + c.gxo;
          double gy = ((double)y - g.y0) / g.dy + c.gyo;
          int bx = ifloor(gx - 0.5), by = ifloor(gy - 0.5);
          int sx, sy;
          if (g.px) {
            int o = pmod(fx - bx, c.uw);
            if (o > 2) continue;
            sx = bx + o;
          } else {
            if (fx < bx || fx > bx + 2) continue;
            sx = fx;
          }
          if (g.py) {
            int o = pmod(fy - by, c.uh);
            if (o > 2) continue;
            sy = by + o;
          } else {
            if (fy < by || fy > by + 2) continue;
            sy = fy;
          }
          double w = bspline(gy - (double)sy) * bspline(gx - (double)sx);
          T vx, vy;
          ld2(a.sval, k, vx, vy);
          double val = (double)(a.comp ? vy : vx);
          wsum = (T)((double)wsum + w);
          msum = (T)((double)msum + w * val);
        }
      }
    }
    T outv;
    bool supported = (double)wsum > 1.0e-12;
    if (supported) outv = msum / wsum;
    else outv = a.fallback[(size_t)fy * c.W + fx];
    // the duplicate endpoint itself never receives weight (kern/pic.py:136-151)
    uns = (!supported || dupc || dupr) ? 1 : 0;
    a.out[(size_t)ty * c.W + tx] = outv;
  }
  uns = wave_sum_i(uns);
  if ((threadIdx.x & 63) == 0 && uns) atomicAdd(a.unsupported, (unsigned long long)uns);
}

// The same gather for a non-periodic box, staged through LDS. A block owns a tile of
// kTileW x kTileH faces and loads once the particles of the cells that can reach it, (kTileW + 3) x
// (kTileH + 2) for u faces and (kTileW + 2) x (kTileH + 3) for v faces (each cell row's slice of the
// sorted arrays is contiguous), with one round of loads. Staging evaluates each
// particle's stencil base and its three B-spline weights per axis once, in FP64 (kern/pic.py:
// 127-147 evaluates them per particle too), and keeps them in the storage type (FP64 weights are
// therefore exact; FP32 weights are rounded once, 2^-24 relative, which keeps two tiles' staging in
// one CU's LDS); a face then takes a candidate's weight as the FP64 product of the two weights its
// offset selects, in FP64 the same product of the same values as evaluating both B-splines for that
// face. A face's candidates are 4 x 3 cells (u) or 3 x 4 (v); four threads share the face, one per
// column (u) or row (v) of them, each sums its three cells' contributions in FP64 in cell and
// particle order, and the four partial sums are added in thread order, so the result is
// deterministic. A face's weight and momentum are rounded
// to the storage type once (the reference rounds after every particle, in particle order). A tile
// whose cells hold more particles than the LDS capacity gathers from global memory instead.
constexpr int kTileW = 32, kTileH = 4, kTileRows = 4;
constexpr int kTileThreads = kTileW * kTileH * kTileRows;
template <class T>
constexpr int tile_cap() {
  // staged particles per tile: six weights and the value in the storage type, the packed base
  // (30 B each in FP32, ~64 KB, two tiles per CU; 58 B in FP64, ~114 KB)
  return sizeof(T) == 4 ? 2150 : 2000;
}
// Stencil base of a staged particle relative to the tile's first face, packed in 16 bits (8 each,
// biased by 8); a base that no face of the tile can use is stored as kNoBase.
constexpr int kBaseBias = 8, kNoBase = 0xffff;
__device__ __forceinline__ int pack_base(int rx, int ry) {
  if (rx < -2 || rx > kTileW + 1 || ry < -2 || ry > kTileH + 1) return kNoBase;
  return (rx + kBaseBias) | ((ry + kBaseBias) << 8);
}

template <class T>
__global__ void __launch_bounds__(kTileThreads) p2g_tile_kernel(P2GArgs<T> a) {
  constexpr int CAP = tile_cap<T>();
  __shared__ T swx[3][CAP], swy[3][CAP];
  __shared__ T sv[CAP];
  __shared__ unsigned short sb[CAP];
  // the four rows' partial sums reuse swx once the gather is done
  double* part = reinterpret_cast<double*>(&swx[0][0]);
  constexpr int NR = kTileH + 3, NC = kTileW + 3;
  // cell_raw[r][lc]: the sorted arrays' start of cell (Y0 - 2 + r, X0 - 2 + lc), clamped into the
  // row's slice of cells X0-2 .. X0+kTileW (clipped to the box); lc = 0 is the slice's first
  // particle, lc = NC one past its last. cell_off: the same positions in the tile's staging order.
  __shared__ int cell_raw[NR][NC + 1];
  __shared__ int cell_off[NR][NC + 1];
  static_assert(NR * (NC + 1) <= kTileThreads, "one thread per cell boundary");
  static_assert(2 * kTileThreads * sizeof(double) <= 3 * CAP * sizeof(T), "the partial sums reuse swx");
  const Comp& c = a.c;
  const PicGeom& g = a.g;
  const int face = threadIdx.y * kTileW + threadIdx.x;
  const int tid = threadIdx.z * (kTileW * kTileH) + face;
  const int X0 = blockIdx.x * kTileW, Y0 = blockIdx.y * kTileH;
  // The cells whose particles can reach the tile's faces. Along the axis where a face sits half a
  // cell off the cell centres (y for u faces, x for v faces) a particle's stencil base is its home
  // cell minus one, (q - 0.5) - 0.5 being exactly q - 1 in FP64, so face f needs cells f-1 .. f+1
  // there; along the other axis the base is the home cell or the one before it: cells f-2 .. f+1.
  // A particle beyond the box, binned to an edge cell, reaches only that edge's faces, which those
  // ranges hold. Staged rows: Y0-1 .. Y0+kTileH (u) or Y0-2 .. (v); columns X0-2 .. X0+kTileW (u)
  // or X0-1 .. (v).
  const bool vc = a.comp != 0;
  const int RY0 = Y0 - (vc ? 2 : 1), RX0 = X0 - (vc ? 1 : 2);
  const int nr = vc ? NR : NR - 1, nc = vc ? NC - 1 : NC;
  const int cx0 = max(RX0, 0), cx1 = min(X0 + kTileW, g.nx - 1);
  // every cell boundary of the tile in one round of loads
  if (tid < NR * (NC + 1)) {
    const int r = tid / (NC + 1), lc = tid - r * (NC + 1);
    const int cy = RY0 + r;
    int v = 0;
    if (r < nr && cy >= 0 && cy < g.ny && cx0 <= cx1) v = a.start[cy * g.nx + min(max(RX0 + lc, cx0), cx1 + 1)];
    cell_raw[r][lc] = v;
  }
  __syncthreads();
  // row r's particles take staging slots row_base[r] ..; every thread forms the NR + 1 bases
  int row_base[NR + 1];
  row_base[0] = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) row_base[r + 1] = row_base[r] + (cell_raw[r][nc] - cell_raw[r][0]);
  if (tid < NR * (NC + 1)) {
    const int r = tid / (NC + 1), lc = tid - r * (NC + 1);
    int rb = 0;
#pragma unroll
    for (int q = 0; q < NR; ++q) rb = q == r ? row_base[q] : rb;
    cell_off[r][lc] = rb + cell_raw[r][lc] - cell_raw[r][0];
  }
  const int total = row_base[NR];
  const bool staged = total <= CAP;
  if (staged) {
    // each thread's slots o = tid + i * kTileThreads, their loads issued together
    constexpr int NIT = (CAP + kTileThreads - 1) / kTileThreads;
    T px[NIT], py[NIT], pv[NIT];
#pragma unroll
    for (int i = 0; i < NIT; ++i) {
      const int o = tid + i * kTileThreads;
      if (o < total) {
        int r = 0;
#pragma unroll
        for (int q = 1; q < NR; ++q) r += o >= row_base[q] ? 1 : 0;
        int rb = 0, first = 0;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          rb = q == r ? row_base[q] : rb;
          first = q == r ? cell_raw[q][0] : first;
        }
        const size_t p = (size_t)first + (o - rb);
        ld2(a.spos, p, px[i], py[i]);
        pv[i] = a.sval[2 * p + (a.comp ? 1 : 0)];
      }
    }
#pragma unroll
    for (int i = 0; i < NIT; ++i) {
      const int o = tid + i * kTileThreads;
      if (o < total) {
        const double gx = ((double)px[i] - g.x0) / g.dx + c.gxo;
        const double gy = ((double)py[i] - g.y0) / g.dy + c.gyo;
        const int bx = ifloor(gx - 0.5), by = ifloor(gy - 0.5);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          swx[q][o] = (T)bspline(gx - (double)(bx + q));
          swy[q][o] = (T)bspline(gy - (double)(by + q));
        }
        sb[o] = (unsigned short)pack_base(bx - X0, by - Y0);
        sv[o] = pv[i];
      }
    }
  }
  __syncthreads();
  // Thread z of a face takes the face's z-th column of candidate cells (u) or row (v), three
  // cells each.
  const int fx = X0 + (int)threadIdx.x, fy = Y0 + (int)threadIdx.y;
  const int z = (int)threadIdx.z;
  double wsum = 0.0, msum = 0.0;
  if (fx < c.W && fy < c.H) {
    for (int j = 0; j < 3; ++j) {
      const int r = (int)threadIdx.y + (vc ? z : j), lc = (int)threadIdx.x + (vc ? j : z);
      const int cx = RX0 + lc, cy = RY0 + r;
      if (cx < 0 || cx >= g.nx || cy < 0 || cy >= g.ny) continue;
      if (staged) {
        const int k0 = cell_off[r][lc], k1 = cell_off[r][lc + 1];
        const int lx = (int)threadIdx.x + kBaseBias, ly = (int)threadIdx.y + kBaseBias;
        for (int k = k0; k < k1; ++k) {
          const int b = sb[k];
          const unsigned ox = (unsigned)(lx - (b & 0xff)), oy = (unsigned)(ly - (b >> 8));
          if (ox > 2u || oy > 2u) continue;
          const double w = (double)swy[oy][k] * (double)swx[ox][k];
          wsum = wsum + w;
          msum = msum + w * (double)sv[k];
        }
      } else {
        const int k0 = a.start[cy * g.nx + cx], k1 = a.start[cy * g.nx + cx + 1];
        for (int k = k0; k < k1; ++k) {
          T x, y;
          ld2(a.spos, k, x, y);
          const double gx = ((double)x - g.x0) / g.dx + c.gxo;
          const double gy = ((double)y - g.y0) / g.dy + c.gyo;
          const int bx = ifloor(gx - 0.5), by = ifloor(gy - 0.5);
          if (fx < bx || fx > bx + 2) continue;
          if (fy < by || fy > by + 2) continue;
          const double w = bspline(gy - (double)fy) * bspline(gx - (double)fx);
          T vx, vy;
          ld2(a.sval, k, vx, vy);
          wsum = wsum + w;
          msum = msum + w * (double)(a.comp ? vy : vx);
        }
      }
    }
  }
  // the four threads' partial sums, added in z order (the staging arrays are free now)
  __syncthreads();
  part[threadIdx.z * (kTileW * kTileH) + face] = wsum;
  part[kTileThreads + threadIdx.z * (kTileW * kTileH) + face] = msum;
  __syncthreads();
  long long uns = 0;
  if (threadIdx.z == 0 && fx < c.W && fy < c.H) {
    double w = part[face], m = part[kTileThreads + face];
#pragma unroll
    for (int q = 1; q < kTileRows; ++q) {
      w = w + part[q * (kTileW * kTileH) + face];
      m = m + part[kTileThreads + q * (kTileW * kTileH) + face];
    }
    const T wt = (T)w, mt = (T)m;
    T outv;
    const bool supported = (double)wt > 1.0e-12;
    if (supported) outv = mt / wt;
    else outv = a.fallback[(size_t)fy * c.W + fx];
    uns = supported ? 0 : 1;
    a.out[(size_t)fy * c.W + fx] = outv;
  }
  uns = wave_sum_i(uns);
  if ((tid & 63) == 0 && uns) atomicAdd(a.unsupported, (unsigned long long)uns);
}

// ------------------------------------------------------------------ G2P and the FLIP update

template <class T>
__global__ void __launch_bounds__(256) g2p_kernel(PicGeom g, int n, const int* __restrict__ idx,
                                                  const T* __restrict__ pos, const T* __restrict__ u,
                                                  const T* __restrict__ v, T* __restrict__ vel) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int i = idx ? idx[k] : k;
  T x, y;
  ld2(pos, i, x, y);
  double a, b, dummy;
  double tx, ty;
  grid_coords(g, (double)x, (double)y, tx, ty);
  gather<T, false>(g, comp_u(g), u, nullptr, tx, ty, a, dummy);
  gather<T, false>(g, comp_v(g), v, nullptr, tx, ty, b, dummy);
  st2(vel, i, (T)a, (T)b);
}

// ------------------------------------------------------------------ cell-centred transfers
// particle_to_grid_kernel and grid_to_particle_kernel (kern/pic.py:22-102): quadratic B-spline
// transfers between particles and a (ny, nx, 2) cell-centred grid, the 3 x 3 stencil wrapped on a
// periodic axis and clamped to the edge cells otherwise, so an edge cell can take several of one
// particle's nine weights.

// P2G onto cells as a gather in the reference's order. Cell (tx, ty) receives only from particles
// whose stencil base lies in ty-2..ty and tx-2..tx (after wrapping or clamping): at most 3 x 3 bins,
// each listing its particles by index. The gather merges those lists by particle index and adds each
// particle's weights for this cell in its 3 x 3 offset order, so every cell sees the sequence of
// additions of the reference's particle-major loop: the FP64 contribution added to the dtype
// accumulator in FP64 and rounded back (kern/pic.py:49-51). Results are bit-identical to it.
template <class T>
__global__ void __launch_bounds__(256) p2c_kernel(PicGeom g, const int* __restrict__ start,
                                                  const int* __restrict__ sidx, const T* __restrict__ spos,
                                                  const T* __restrict__ sval, double fsx, double fsy,
                                                  T* __restrict__ grid) {
  const int tx = blockIdx.x * 256 + threadIdx.x;
  const int ty = blockIdx.y;
  if (tx >= g.nx) return;
  const int bw = base_bins(g.nx, g.px);
  // the bins of bases ty-2..ty and tx-2..tx, each taken once (they repeat at a clamped edge and on
  // a periodic axis of fewer than three cells)
  int rows[3], cols[3];
  bool rnew[3], cnew[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    rows[k] = base_bin(ty - 2 + k, g.ny, g.py);
    cols[k] = base_bin(tx - 2 + k, g.nx, g.px);
    rnew[k] = cnew[k] = true;
#pragma unroll
    for (int j = 0; j < k; ++j) {
      rnew[k] = rnew[k] && rows[j] != rows[k];
      cnew[k] = cnew[k] && cols[j] != cols[k];
    }
  }
  // cursor, end and head particle index of each candidate bin (INT_MAX once drained)
  int cur[9], end[9], head[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const int r = j / 3, c = j % 3;
    cur[j] = end[j] = 0;
    head[j] = INT_MAX;
    if (rnew[r] && cnew[c]) {
      const int bin = rows[r] * bw + cols[c];
      cur[j] = start[bin];
      end[j] = start[bin + 1];
      if (cur[j] < end[j]) head[j] = sidx[cur[j]];
    }
  }
  T w = (T)0, mx = (T)0, my = (T)0;
  for (;;) {
    int best = -1, bi = INT_MAX;
#pragma unroll
    for (int j = 0; j < 9; ++j)
      if (head[j] < bi) {
        bi = head[j];
        best = j;
      }
    if (best < 0) break;
    int k = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j)
      if (j == best) {
        k = cur[j]++;
        head[j] = cur[j] < end[j] ? sidx[cur[j]] : INT_MAX;
      }
    T x, y, vx, vy;
    ld2(spos, k, x, y);
    ld2(sval, k, vx, vy);
    const double gx = cell_coord((double)x, g.x0, g.dx), gy = cell_coord((double)y, g.y0, g.dy);
    const int bx = ifloor(gx - 0.5), by = ifloor(gy - 0.5);
    for (int oy = 0; oy < 3; ++oy) {
      const int sy = by + oy;
      if (cell_target(sy, g.ny, g.py) != ty) continue;
      const double wy = bspline(gy - (double)sy);
      for (int ox = 0; ox < 3; ++ox) {
        const int sx = bx + ox;
        if (cell_target(sx, g.nx, g.px) != tx) continue;
        const double wt = wy * bspline(gx - (double)sx);
        w = (T)((double)w + wt);
        mx = (T)((double)mx + wt * (double)vx);
        my = (T)((double)my + wt * (double)vy);
      }
    }
  }
  T ox, oy;
  if ((double)w > 1.0e-12) {
    // the dtype quotient: an FP64 quotient of FP32 operands rounds to the correctly rounded FP32 one
    ox = (T)((double)mx / (double)w);
    oy = (T)((double)my / (double)w);
  } else {
    ox = (T)fsx;
    oy = (T)fsy;
  }
  st2(grid, (size_t)ty * g.nx + tx, ox, oy);
}

// grid_to_particle_kernel (kern/pic.py:67-102): FP64 sums in the 3 x 3 order, scaled by
// 1 / max(weight sum, 1e-12), rounded to the dtype.
template <class T>
__global__ void __launch_bounds__(256) c2p_kernel(PicGeom g, int n, const T* __restrict__ grid,
                                                  const T* __restrict__ pos, T* __restrict__ vel) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  T x, y;
  ld2(pos, k, x, y);
  const double gx = cell_coord((double)x, g.x0, g.dx), gy = cell_coord((double)y, g.y0, g.dy);
  const int bx = ifloor(gx - 0.5), by = ifloor(gy - 0.5);
  double vx = 0.0, vy = 0.0, ws = 0.0;
  for (int oy = 0; oy < 3; ++oy) {
    const int sy = by + oy;
    const int ty = cell_target(sy, g.ny, g.py);
    const double wy = bspline(gy - (double)sy);
    for (int ox = 0; ox < 3; ++ox) {
      const int sx = bx + ox;
      const int tx = cell_target(sx, g.nx, g.px);
      const double w = wy * bspline(gx - (double)sx);
      T a, b;
      ld2(grid, (size_t)ty * g.nx + tx, a, b);
      vx += w * (double)a;
      vy += w * (double)b;
      ws += w;
    }
  }
  const double inv = 1.0 / fmax(ws, 1.0e-12);
  st2(vel, k, (T)(vx * inv), (T)(vy * inv));
}


// ------------------------------------------------------------------ collisions

// ω x r wall velocity (fb/core/geometry.py:131-138): the FP64 angular velocity times the
// dtype offset from the pivot, rounded to the dtype.
template <class T>
__device__ __forceinline__ void wall_t(const NacaDev& g, double omega, T x, T y, T& wx, T& wy) {
  T rx = x - (T)g.pivot_x, ry = y - (T)g.pivot_y;
  wx = (T)((-omega) * (double)ry);
  wy = (T)(omega * (double)rx);
}

// Remove the inward wall-relative normal velocity (fb/solvers/pic_flip.py:589-596, 612-617).
template <class T>
__device__ __forceinline__ void reflect_t(T& vx, T& vy, T wx, T wy, T nx, T ny) {
  T rx = vx - wx, ry = vy - wy;
  T inward = rx * nx + ry * ny;
  if (inward < (T)0) {
    rx = rx - inward * nx;
    ry = ry - inward * ny;
  }
  vx = wx + rx;
  vy = wy + ry;
}

// _resolve_particle_collisions (fb/solvers/pic_flip.py:604-618).
template <class T>
__global__ void __launch_bounds__(256) collide_kernel(int n, T* __restrict__ pos, T* __restrict__ vel, NacaDev g,
                                                      double omega) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  T x, y;
  ld2(pos, i, x, y);
  T d = naca_sdf_t<T>(g, x, y);
  if (!(d <= (T)0)) return;
  T nx, ny;
  naca_normal_t<T>(g, x, y, nx, ny);
  T s = d - (T)1.0e-4;
  x = x - s * nx;
  y = y - s * ny;
  st2(pos, i, x, y);
  T wx, wy, vx, vy;
  wall_t(g, omega, x, y, wx, wy);
  ld2(vel, i, vx, vy);
  reflect_t(vx, vy, wx, wy, nx, ny);
  st2(vel, i, vx, vy);
}

// ------------------------------------------------------------------ FLIP update and advection

struct FlipArgs {
  PicGeom g;
  double blend, dt;
  double pivot_x, pivot_y;
  double reach;  // maximum_radius + collision margin
  unsigned long long* travel_key;
};

// Blend (fb/solvers/pic_flip.py:817-831), then the midpoint step of _advect_particles
// (:489-494) with the PIC velocity, and the swept-collision candidate test (:533-550).
template <class T>
__global__ void __launch_bounds__(256) flip_advect_kernel(FlipArgs a, int n, T* __restrict__ pos,
                                                          T* __restrict__ vel, T* __restrict__ start,
                                                          uint8_t* __restrict__ cand, const T* __restrict__ qu,
                                                          const T* __restrict__ qv, const T* __restrict__ pu,
                                                          const T* __restrict__ pv) {
  int i = blockIdx.x * 256 + threadIdx.x;
  double travel = 0.0;
  if (i < n) {
    const PicGeom& g = a.g;
    const Comp cu = comp_u(g), cv = comp_v(g);
    T x, y, vx, vy;
    ld2(pos, i, x, y);
    ld2(vel, i, vx, vy);
    double qa, da, qb, db, tx, ty;
    grid_coords(g, (double)x, (double)y, tx, ty);
    gather<T, true>(g, cu, qu, pu, tx, ty, qa, da);
    gather<T, true>(g, cv, qv, pv, tx, ty, qb, db);
    const T picx = (T)qa, picy = (T)qb, dlx = (T)da, dly = (T)db;
    const T omb = (T)(1.0 - a.blend), b = (T)a.blend;
    vx = omb * picx + b * (vx + dlx);
    vy = omb * picy + b * (vy + dly);
    st2(vel, i, vx, vy);
    // midpoint = positions + 0.5 * dt * velocity_0; positions += dt * velocity_mid
    const T h = (T)(0.5 * a.dt), dt = (T)a.dt;
    T mx = x + h * picx, my = y + h * picy;
    double ma, mb, dummy;
    grid_coords(g, (double)mx, (double)my, tx, ty);
    gather<T, false>(g, cu, qu, nullptr, tx, ty, ma, dummy);
    gather<T, false>(g, cv, qv, nullptr, tx, ty, mb, dummy);
    T nx = x + dt * (T)ma, ny = y + dt * (T)mb;
    st2(start, i, x, y);
    st2(pos, i, nx, ny);
    // candidate: the path passes within maximum_radius + margin of the pivot
    T px = nx - x, py = ny - y;
    T pvx = (T)a.pivot_x, pvy = (T)a.pivot_y;
    T l2 = px * px + py * py;
    T l2c = l2 > (T)1.0e-12 ? l2 : (T)1.0e-12;
    T proj = ((pvx - x) * px + (pvy - y) * py) / l2c;
    proj = proj < (T)0 ? (T)0 : (proj > (T)1 ? (T)1 : proj);
    T cx = x + proj * px - pvx, cy = y + proj * py - pvy;
    T dist = sqrt(cx * cx + cy * cy);
    bool c = dist <= (T)a.reach;
    cand[i] = c ? 1 : 0;
    if (c) travel = (double)sqrt(px * px + py * py);
  }
  travel = wave_max(travel);
  if ((threadIdx.x & 63) == 0 && travel > 0.0) key_max(a.travel_key, dkey(travel));
  // a key of 0 (no candidate moved) reads as zero travel
}

struct SweptArgs {
  NacaDev g;           // foil with the pivot; cos/sin per sample from the table
  const double* table;  // (cos, sin) per sample: counts 2..16, samples 1..count
  double omega;
  double wall_travel;  // |deg2rad(angle change)| * maximum_radius
  double spacing;      // 0.1 * min(dx, dy)
  double margin;       // 0.05 * min(dx, dy)
  const unsigned long long* travel_key;
  unsigned long long* hits;
};

// _resolve_swept_particle_collisions (fb/solvers/pic_flip.py:527-602) for the candidates.
template <class T>
__global__ void __launch_bounds__(256) swept_kernel(SweptArgs a, int n, T* __restrict__ pos, T* __restrict__ vel,
                                                    const T* __restrict__ start, const uint8_t* __restrict__ cand) {
  int i = blockIdx.x * 256 + threadIdx.x;
  long long hit = 0;
  if (i < n && cand[i]) {
    const unsigned long long key = *a.travel_key;
    double travel = key ? dkey_inv(key) : 0.0;
    double q = ceil((travel + a.wall_travel) / fmax(a.spacing, 1.0e-12));
    int count = (int)fmin(fmax(q, 2.0), 16.0);
    const int off = (count - 1) * count / 2 - 1;
    T sx, sy, ex, ey;
    ld2(start, i, sx, sy);
    ld2(pos, i, ex, ey);
    const T px = ex - sx, py = ey - sy;
    const T margin = (T)a.margin;
    for (int s = 1; s <= count; ++s) {
      const double frac = (double)s / (double)count;
      NacaDev gs = a.g;
      gs.cos_a = a.table[2 * (off + s - 1)];
      gs.sin_a = a.table[2 * (off + s - 1) + 1];
      const T f = (T)frac;
      T qx = sx + f * px, qy = sy + f * py;
      T d = naca_sdf_t<T>(gs, qx, qy);
      if (!(d <= margin)) continue;
      T nx, ny;
      naca_normal_t<T>(gs, qx, qy, nx, ny);
      T e = d - margin;
      T rx = qx - e * nx, ry = qy - e * ny;
      st2(pos, i, rx, ry);
      T wx, wy, vx, vy;
      wall_t(a.g, a.omega, rx, ry, wx, wy);
      ld2(vel, i, vx, vy);
      reflect_t(vx, vy, wx, wy, nx, ny);
      st2(vel, i, vx, vy);
      hit = 1;
      break;
    }
  }
  hit = wave_sum_i(hit);
  if ((threadIdx.x & 63) == 0 && hit) atomicAdd(a.hits, (unsigned long long)hit);
}

// np.mod for floats of the dtype (Python sign convention).
template <class T>
__device__ __forceinline__ T pymod_t(T a, T b) {
  T m = fmod(a, b);
  if (m != (T)0) {
    if ((b < (T)0) != (m < (T)0)) m += b;
  } else {
    m = copysign((T)0, b);
  }
  return m;
}

// Periodic wrap and escape flags (fb/solvers/pic_flip.py:499-511).
template <class T>
__global__ void __launch_bounds__(256) wrap_escape_kernel(PicGeom g, int n, T* __restrict__ pos,
                                                          int* __restrict__ flag) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  T x, y;
  ld2(pos, i, x, y);
  const T x0 = (T)g.x0, x1 = (T)g.x1, y0 = (T)g.y0, y1 = (T)g.y1;
  if (g.px) x = x0 + pymod_t<T>(x - x0, (T)(g.x1 - g.x0));
  if (g.py) y = y0 + pymod_t<T>(y - y0, (T)(g.y1 - g.y0));
  if (g.px || g.py) st2(pos, i, x, y);
  bool esc = false;
  if (!g.px) esc = esc || (x < x0) || (x > x1);
  if (!g.py) esc = esc || (y < y0) || (y > y1);
  flag[i] = esc ? 1 : 0;
}

// Escape respawn (fb/solvers/pic_flip.py:512-524): escaped particle r (in index order) takes
// draw r for x and draw count + r for y.
template <class T>
__global__ void __launch_bounds__(256) respawn_kernel(PicGeom g, int n, const int* __restrict__ flag,
                                                      const int* __restrict__ rank, const int* __restrict__ total,
                                                      const Pcg* __restrict__ rng, double fsx, double fsy,
                                                      T* __restrict__ pos, T* __restrict__ vel) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const Pcg s = *rng;
  const unsigned long long r = (unsigned long long)rank[i];
  const unsigned long long c = (unsigned long long)*total;
  float xs;
  if (g.px) {
    xs = pcg_uniform(pcg_draw(s, r), g.x0, g.x1);
  } else {
    xs = (float)g.x0 + pcg_uniform(pcg_draw(s, r), 0.0, 0.5 * g.dx);
  }
  float ys = pcg_uniform(pcg_draw(s, c + r), g.y0, g.y1);
  st2(pos, i, (T)xs, (T)ys);
  st2(vel, i, (T)fsx, (T)fsy);
}

// Advance the stream state by 2 * total (or by `fixed` draws when total is null).
__global__ void rng_advance_kernel(Pcg* rng, const int* total, unsigned long long fixed) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  unsigned long long d = total ? 2ull * (unsigned long long)*total : fixed;
  rng->state = pcg_advance(rng->state, rng->inc, d);
}

// ------------------------------------------------------------------ seeding

__global__ void __launch_bounds__(256) fluid_flag_kernel(int ncell, const uint8_t* __restrict__ solid,
                                                         int* __restrict__ flag) {
  int c = blockIdx.x * 256 + threadIdx.x;
  if (c < ncell) flag[c] = solid[c] ? 0 : 1;
}

__global__ void __launch_bounds__(256) fluid_list_kernel(int ncell, const uint8_t* __restrict__ solid,
                                                         const int* __restrict__ rank, int* __restrict__ list) {
  int c = blockIdx.x * 256 + threadIdx.x;
  if (c < ncell && !solid[c]) list[rank[c]] = c;
}

// _seed_particles (fb/solvers/pic_flip.py:317-333): `per_cell` particles per fluid cell in
// row-major order, jitter draws 2k and 2k+1 for particle k.
template <class T>
__global__ void __launch_bounds__(256) seed_kernel(PicGeom g, int n, int per_cell, const int* __restrict__ list,
                                                   const Pcg* __restrict__ rng, T* __restrict__ pos) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const Pcg s = *rng;
  const int c = list[k / per_cell];
  const int cx = c % g.nx, cy = c / g.nx;
  const T jx = (T)pcg_draw(s, 2ull * (unsigned long long)k);
  const T jy = (T)pcg_draw(s, 2ull * (unsigned long long)k + 1ull);
  const double tx = ((double)cx + 0.1) + (double)((T)0.8 * jx);
  const double ty = ((double)cy + 0.1) + (double)((T)0.8 * jy);
  st2(pos, k, (T)(g.x0 + tx * g.dx), (T)(g.y0 + ty * g.dy));
}

// ------------------------------------------------------------------ statistics

// Reference cell ids (fb/core/particle_population.py:19-31): dtype arithmetic, clipped.
template <class T>
__device__ __forceinline__ int ref_cell(const PicGeom& g, T x, T y) {
  T rx = floor((x - (T)g.x0) / (T)g.dx);
  T ry = floor((y - (T)g.y0) / (T)g.dy);
  double fx = fmin(fmax((double)rx, 0.0), (double)(g.nx - 1));
  double fy = fmin(fmax((double)ry, 0.0), (double)(g.ny - 1));
  if (!(fx == fx)) fx = 0.0;
  if (!(fy == fy)) fy = 0.0;
  return (int)fy * g.nx + (int)fx;
}

// Partial slots of the statistics (folded into the keys): [0] max particle speed, [1] particles
// inside the foil, [2] non-finite particle values, [3] max grid speed, [4] non-finite grid values,
// [5] fluid cells, [6] empty fluid cells, [7] fluid cells with fewer than 4 particles.
template <class T>
__global__ void __launch_bounds__(256) pic_particle_stats_kernel(PicGeom g, int n, const T* __restrict__ pos,
                                                                 const T* __restrict__ vel, NacaDev foil,
                                                                 int* __restrict__ counts,
                                                                 unsigned long long* __restrict__ part) {
  __shared__ unsigned long long lds[4];
  unsigned long long sp = 0, inside = 0, bad = 0;
  // whole waves step together so wave_run_add sees every lane
  for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
    const int i = base + threadIdx.x;
    const bool act = i < n;
    int c = 0;
    if (act) {
      T x, y, vx, vy;
      ld2(pos, i, x, y);
      ld2(vel, i, vx, vy);
      if (!isfinite((double)x) || !isfinite((double)y)) bad++;
      if (!isfinite((double)vx) || !isfinite((double)vy)) bad++;
      const double s = (double)sqrt(vx * vx + vy * vy);
      if (s == s) {
        const unsigned long long k = dkey(s);
        sp = k > sp ? k : sp;
      }
      if (naca_sdf_t<T>(foil, x, y) <= (T)0) inside++;
      c = ref_cell(g, x, y);
    }
    wave_run_add(counts, c, act);
  }
  put_max(part, 0, sp, lds);
  put_sum(part, 1, inside, lds);
  put_sum(part, 2, bad, lds);
}

template <class T>
__global__ void __launch_bounds__(256) pic_cell_stats_kernel(int ncell, const T* __restrict__ gu,
                                                             const T* __restrict__ gv,
                                                             const uint8_t* __restrict__ solid,
                                                             const int* __restrict__ counts,
                                                             unsigned long long* __restrict__ part) {
  __shared__ unsigned long long lds[4];
  unsigned long long sp = 0, bad = 0, fluid = 0, empty = 0, under = 0;
  for (int c = blockIdx.x * 256 + threadIdx.x; c < ncell; c += gridDim.x * 256) {
    T a = gu[c], b = gv[c];
    if (!isfinite((double)a) || !isfinite((double)b)) bad++;
    const double s = (double)sqrt(a * a + b * b);
    if (s == s) {
      const unsigned long long k = dkey(s);
      sp = k > sp ? k : sp;
    }
    if (!solid[c]) {
      fluid++;
      empty += counts[c] == 0;
      under += counts[c] < 4;
    }
  }
  put_max(part, 3, sp, lds);
  put_sum(part, 4, bad, lds);
  put_sum(part, 5, fluid, lds);
  put_sum(part, 6, empty, lds);
  put_sum(part, 7, under, lds);
}

__global__ void fill_u64_kernel(unsigned long long* a, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0ull;
}

}  // namespace fb

// ====================================================================== host side

namespace fb {

struct SortBufs {
  int ncell = 0, cap = 0;
  int* cellof = nullptr;
  int* sidx = nullptr;
  int* counts = nullptr;  // ncell + 1
  int* cstart = nullptr;  // ncell + 1
  int* cursor = nullptr;  // ncell
  int* bsum = nullptr;
  int* total = nullptr;  // 4 ints
  void* spos = nullptr;
  void* sval = nullptr;
};

static int alloc_sort(SortBufs& b, int ncell, int cap, size_t sz, hipStream_t s, bool async) {
  b.ncell = ncell;
  b.cap = cap;
  const size_t nb = (size_t)(std::max(ncell, cap) + 1023) / 1024 + 2;
  auto A = [&](void** p, size_t bytes) -> hipError_t {
    return async ? hipMallocAsync(p, std::max<size_t>(bytes, 16), s) : dev_alloc(p, std::max<size_t>(bytes, 16));
  };
  hipError_t e = hipSuccess;
  void** ptrs[] = {(void**)&b.cellof, (void**)&b.sidx, (void**)&b.counts, (void**)&b.cstart, (void**)&b.cursor,
                   (void**)&b.bsum, (void**)&b.total, &b.spos, &b.sval};
  size_t bytes[] = {(size_t)cap * 4, (size_t)cap * 4, ((size_t)ncell + 1) * 4, ((size_t)ncell + 1) * 4,
                    (size_t)ncell * 4, nb * 4, 16, (size_t)cap * 2 * sz, (size_t)cap * 2 * sz};
  for (int k = 0; k < 9 && e == hipSuccess; ++k) e = A(ptrs[k], bytes[k]);
  if (e != hipSuccess) {
    set_error("device allocation failed: %s", hipGetErrorString(e));
    return FB_ERR_NOMEM;
  }
  return FB_OK;
}

static void free_sort(SortBufs& b, hipStream_t s, bool async) {
  for (void* p : {(void*)b.cellof, (void*)b.sidx, (void*)b.counts, (void*)b.cstart, (void*)b.cursor,
                  (void*)b.bsum, (void*)b.total, b.spos, b.sval})
    if (p) (void)(async ? hipFreeAsync(p, s) : dev_free(p));
  b = SortBufs{};
}

// Exclusive scan of n ints: out[0..n) prefixes, out[n] = total (also written to *total).
static int scan_ints(hipStream_t s, const int* in, int* out, int n, int* bsum, int* total) {
  const int nb = std::max(1, (n + 1023) / 1024);
  hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(256), 0, s, in, out, n, bsum);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(256), 0, s, bsum, nb, total);
  hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(256), 0, s, out, n, bsum, total);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

static inline unsigned blocks(long long n) { return (unsigned)std::max<long long>(1, (n + 255) / 256); }

// Counting sort of the particles by bin (b.ncell bins: home cells, or with BASE the cell-centred stencil
// bases), each bin's list ordered by particle index, with the positions and values copied in that order.
template <class T, bool BASE = false>
static int sort_by_cell(hipStream_t s, const PicGeom& g, int n, const T* pos, const T* vel, SortBufs& b) {
  const int nc = b.ncell;
  hipLaunchKernelGGL(fill_int_kernel, dim3(blocks(nc + 1)), dim3(256), 0, s, b.counts, nc + 1, 0);
  hipLaunchKernelGGL(fill_int_kernel, dim3(blocks(nc)), dim3(256), 0, s, b.cursor, nc, 0);
  if (n > 0)
    hipLaunchKernelGGL((bin_count_kernel<T, BASE>), dim3(blocks(n)), dim3(256), 0, s, g, n, pos, b.cellof,
                       b.counts);
  FB_HIP(hipGetLastError());
  int rc = scan_ints(s, b.counts, b.cstart, nc, b.bsum, b.total);
  if (rc) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(bin_scatter_kernel, dim3(blocks(n)), dim3(256), 0, s, n, b.cellof, b.cstart, b.cursor, b.sidx);
    hipLaunchKernelGGL(cell_sort_kernel, dim3(blocks(nc)), dim3(256), 0, s, nc, b.cstart, b.sidx);
    hipLaunchKernelGGL(sorted_copy_kernel<T>, dim3(blocks(n)), dim3(256), 0, s, n, b.sidx, pos, vel, (T*)b.spos,
                       (T*)b.sval);
  }
  FB_HIP(hipGetLastError());
  return FB_OK;
}

template <class T>
static int launch_p2g(hipStream_t s, const PicGeom& g, const SortBufs& b, const T* fu, const T* fv, T* tu, T* tv,
                      unsigned long long* unsupported) {
  for (int c = 0; c < 2; ++c) {
    P2GArgs<T> a;
    a.g = g;
    a.c = c ? comp_v(g) : comp_u(g);
    a.comp = c;
    a.start = b.cstart;
    a.spos = (const T*)b.spos;
    a.sval = (const T*)b.sval;
    a.fallback = c ? fv : fu;
    a.out = c ? tv : tu;
    a.unsupported = unsupported;
    if (!g.px && !g.py) {
      dim3 grid((a.c.W + kTileW - 1) / kTileW, (a.c.H + kTileH - 1) / kTileH);
      hipLaunchKernelGGL(p2g_tile_kernel<T>, grid, dim3(kTileW, kTileH, kTileRows), 0, s, a);
    } else {
      dim3 grid((a.c.W + 255) / 256, a.c.H);
      hipLaunchKernelGGL(p2g_kernel<T>, grid, dim3(256), 0, s, a);
    }
  }
  FB_HIP(hipGetLastError());
  return FB_OK;
}

template <class T>
static int launch_g2p(hipStream_t s, const PicGeom& g, int n, const int* idx, const T* pos, const T* u, const T* v,
                      T* vel) {
  if (n <= 0) return FB_OK;
  hipLaunchKernelGGL(g2p_kernel<T>, dim3(blocks(n)), dim3(256), 0, s, g, n, idx, pos, u, v, vel);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

static void fill_keys(hipStream_t s, unsigned long long* k, int n) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, k, n);
}

// Number of sweep-table entries: sample counts 2..16, samples 1..count.
constexpr int kSweepEntries = 135;

struct Pic {
  Ctx* ctx = nullptr;
  MacGrid* M = nullptr;
  PicGeom g{};
  fb_foil foil{};
  int prec = FB_F64;
  size_t sz = 8;
  int n = 0, cap = 0;
  void *pos = nullptr, *vel = nullptr, *start = nullptr;
  uint8_t* cand = nullptr;
  int* flag = nullptr;  // cap + 1
  int* rank = nullptr;  // cap + 1
  int* flist = nullptr;  // fluid cell list (ncell)
  SortBufs sb;
  void *GU = nullptr, *GV = nullptr;  // grid velocity planes
  void *FU = nullptr, *FV = nullptr;  // fallback faces
  void *TU = nullptr, *TV = nullptr;  // transferred (pre-projection) faces
  Pcg* rng = nullptr;
  unsigned long long* keys = nullptr;  // 16
  unsigned long long* host_keys = nullptr;
  int* cnt = nullptr;  // per-cell counts for statistics (ncell)
  double* table = nullptr;
  // checkpoint
  void *ck_pos = nullptr, *ck_vel = nullptr, *ck_gu = nullptr, *ck_gv = nullptr;
  Pcg* ck_rng = nullptr;
  int ck_n = 0;
  bool ck_valid = false;
  bool grid_valid = false;
};

static int pic_alloc_particles(Pic& P, int cap) {
  P.ck_valid = false;
  for (void* p : {P.pos, P.vel, P.start, (void*)P.cand, (void*)P.flag, (void*)P.rank, P.ck_pos, P.ck_vel})
    if (p) (void)dev_free(p);
  P.pos = P.vel = P.start = P.ck_pos = P.ck_vel = nullptr;
  P.cand = nullptr;
  P.flag = P.rank = nullptr;
  free_sort(P.sb, P.ctx->stream, false);
  P.cap = cap;
  const size_t pb = (size_t)std::max(cap, 1) * 2 * P.sz;
  hipError_t e = hipSuccess;
  auto A = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = dev_alloc(p, bytes);
  };
  A(&P.pos, pb);
  A(&P.vel, pb);
  A(&P.start, pb);
  A(&P.ck_pos, pb);
  A(&P.ck_vel, pb);
  A((void**)&P.cand, (size_t)std::max(cap, 1));
  A((void**)&P.flag, ((size_t)cap + 1) * 4);
  A((void**)&P.rank, ((size_t)cap + 1) * 4);
  if (e != hipSuccess) {
    set_error("device allocation failed: %s", hipGetErrorString(e));
    return FB_ERR_NOMEM;
  }
  return alloc_sort(P.sb, P.g.nx * P.g.ny, std::max(cap, 1), P.sz, P.ctx->stream, false);
}

static void pic_free(Pic* P) {
  if (!P) return;
  for (void* p : {P->pos, P->vel, P->start, (void*)P->cand, (void*)P->flag, (void*)P->rank, (void*)P->flist,
                  P->GU, P->GV, P->FU, P->FV, P->TU, P->TV, (void*)P->rng, (void*)P->keys, (void*)P->cnt,
                  (void*)P->table, P->ck_pos, P->ck_vel, P->ck_gu, P->ck_gv, (void*)P->ck_rng})
    if (p) (void)dev_free(p);
  free_sort(P->sb, nullptr, false);
  if (P->host_keys) (void)host_free(P->host_keys);
  if (P->M) free_mac(P->M);
  delete P;
}

static int pic_read_keys(Pic& P, int n) {
  FB_HIP(hipMemcpyAsync(P.host_keys, P.keys, n * sizeof(unsigned long long), hipMemcpyDeviceToHost, P.ctx->stream));
  FB_HIP(hipStreamSynchronize(P.ctx->stream));
  return FB_OK;
}

// Fallback faces from the grid velocity (fb/solvers/pic_flip.py:779-781).
template <class T>
static int pic_fallback(Pic& P) {
  return launch_cell_to_faces<T>(*P.M, (const T*)P.GU, (const T*)P.GV, (T*)P.FU, (T*)P.FV);
}

template <class T>
static int pic_collide(Pic& P, const fb_pose& pose, double omega) {
  if (P.n <= 0) return FB_OK;
  hipLaunchKernelGGL(collide_kernel<T>, dim3(blocks(P.n)), dim3(256), 0, P.ctx->stream, P.n, (T*)P.pos, (T*)P.vel,
                     make_naca(P.foil, pose), omega);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

// P2G with the grid velocity as fallback; the unsupported face count lands in keys[slot].
template <class T>
static int pic_p2g(Pic& P, int slot) {
  hipStream_t s = P.ctx->stream;
  int rc;
  if ((rc = pic_fallback<T>(P))) return rc;
  if ((rc = sort_by_cell<T>(s, P.g, P.n, (const T*)P.pos, (const T*)P.vel, P.sb))) return rc;
  return launch_p2g<T>(s, P.g, P.sb, (const T*)P.FU, (const T*)P.FV, (T*)P.TU, (T*)P.TV, P.keys + slot);
}

template <class T>
static int pic_seed(Pic& P, int per_cell, int64_t* count) {
  hipStream_t s = P.ctx->stream;
  MacGrid& M = *P.M;
  const int nc = P.g.nx * P.g.ny;
  int rc;
  // fluid cells in row-major order
  hipLaunchKernelGGL(fluid_flag_kernel, dim3(blocks(nc)), dim3(256), 0, s, nc, M.solid, P.sb.counts);
  if ((rc = scan_ints(s, P.sb.counts, P.sb.cstart, nc, P.sb.bsum, P.sb.total))) return rc;
  hipLaunchKernelGGL(fluid_list_kernel, dim3(blocks(nc)), dim3(256), 0, s, nc, M.solid, P.sb.cstart, P.flist);
  FB_HIP(hipGetLastError());
  int fluid = 0;
  FB_HIP(hipMemcpyAsync(&fluid, P.sb.total, sizeof(int), hipMemcpyDeviceToHost, s));
  FB_HIP(hipStreamSynchronize(s));
  FB_REQUIRE(fluid > 0, FB_ERR_ARG, "particle seeding requires at least one fluid cell");
  const long long n = (long long)fluid * per_cell;
  FB_REQUIRE(n < (1ll << 30), FB_ERR_ARG, "particle count %lld exceeds the supported maximum", n);
  if (n > P.cap) {
    if ((rc = pic_alloc_particles(P, (int)n))) return rc;
  }
  P.n = (int)n;
  hipLaunchKernelGGL(seed_kernel<T>, dim3(blocks(n)), dim3(256), 0, s, P.g, P.n, per_cell, P.flist, P.rng,
                     (T*)P.pos);
  hipLaunchKernelGGL(rng_advance_kernel, dim3(1), dim3(64), 0, s, P.rng, (const int*)nullptr,
                     2ull * (unsigned long long)n);
  FB_HIP(hipGetLastError());
  // velocities from the grid (fb/solvers/pic_flip.py:333, 379-387)
  if ((rc = pic_fallback<T>(P))) return rc;
  if ((rc = launch_g2p<T>(s, P.g, P.n, nullptr, (const T*)P.pos, (const T*)P.FU, (const T*)P.FV, (T*)P.vel)))
    return rc;
  *count = n;
  FB_HIP(hipStreamSynchronize(s));
  return FB_OK;
}

template <class T>
static int pic_substep(Pic& P, const fb_pic_params& p, const fb_pose& pose, const double* table,
                       fb_pic_report* rep) {
  hipStream_t s = P.ctx->stream;
  MacGrid& M = *P.M;
  int rc;
  // sweep table first: the copy is staged before the host waits inside this substep
  FB_HIP(hipMemcpyAsync(P.table, table, 2 * kSweepEntries * sizeof(double), hipMemcpyHostToDevice, s));
  // keys: [0] unsupported faces [1] travel [2] swept hits
  fill_keys(s, P.keys, 4);
  M.last_stats_valid = false;
  if (!M.geom_valid || pose.angle_degrees != M.angle) {
    if ((rc = mac_geometry(M, pose))) return rc;
  }
  if ((rc = pic_collide<T>(P, pose, p.wall_omega_rad))) return rc;
  if ((rc = pic_p2g<T>(P, 0))) return rc;
  // implicit viscosity on the faces (fb/core/grid.py:278-303)
  const int nx = P.g.nx, ny = P.g.ny;
  T* res[2];
  if ((rc = diffuse<T>(M, (const T*)P.TU, (const T*)P.TV, nx + 1 - P.g.px, ny, nx + 1, nx, ny + 1 - P.g.py, nx,
                       p.viscosity, p.dt, p.tolerance, p.max_iterations, res, &rep->grid)))
    return rc;
  if (!rep->grid.viscosity_converged) {
    rep->grid.status = FB_SF_VISCOSITY_NOT_CONVERGED;
    return FB_OK;
  }
  if ((rc = launch_dup_copy<T>(M, res[0], ny, nx + 1, nx + 1, P.g.px, 0))) return rc;
  if ((rc = launch_dup_copy<T>(M, res[1], ny + 1, nx, nx, 0, P.g.py))) return rc;
  FB_HIP(hipMemcpyAsync(M.U[M.cur], res[0], M.NU * M.sz, hipMemcpyDeviceToDevice, s));
  FB_HIP(hipMemcpyAsync(M.V[M.cur], res[1], M.NV * M.sz, hipMemcpyDeviceToDevice, s));
  // projection (fb/solvers/pic_flip.py:436-479): grid.project_faces without the Stable Fluids CFL guard
  fb_sf_params sp{};
  sp.dt = p.dt;
  sp.viscosity = p.viscosity;
  sp.tolerance = p.tolerance;
  sp.max_iterations = p.max_iterations;
  sp.freestream_x = p.freestream_x;
  sp.freestream_y = p.freestream_y;
  sp.wall_omega_rad = p.wall_omega_rad;
  sp.cfl_limit = INFINITY;
  sp.channel_walls = p.channel_walls;
  sp.mixed_poisson = p.mixed_poisson;
  if ((rc = project<T>(M, sp, p.dt, &rep->grid))) return rc;
  if (rep->grid.status) return FB_OK;
  const T* qu = (const T*)M.U[M.cur];
  const T* qv = (const T*)M.V[M.cur];
  if ((rc = launch_faces_to_cell<T>(M, qu, qv, (T*)P.GU, (T*)P.GV))) return rc;
  if (P.n > 0) {
    FlipArgs fa;
    fa.g = P.g;
    fa.blend = p.blend;
    fa.dt = p.dt;
    fa.pivot_x = P.foil.pivot_x;
    fa.pivot_y = P.foil.pivot_y;
    const double margin = 0.05 * std::min(P.g.dx, P.g.dy);
    fa.reach = p.maximum_radius + margin;
    fa.travel_key = P.keys + 1;
    hipLaunchKernelGGL(flip_advect_kernel<T>, dim3(blocks(P.n)), dim3(256), 0, s, fa, P.n, (T*)P.pos, (T*)P.vel,
                       (T*)P.start, P.cand, qu, qv, (const T*)P.TU, (const T*)P.TV);
    SweptArgs sa;
    sa.g = make_naca(P.foil, pose);
    sa.table = P.table;
    sa.omega = p.wall_omega_rad;
    sa.wall_travel = p.wall_travel;
    sa.spacing = 0.1 * std::min(P.g.dx, P.g.dy);
    sa.margin = margin;
    sa.travel_key = P.keys + 1;
    sa.hits = P.keys + 2;
    hipLaunchKernelGGL(swept_kernel<T>, dim3(blocks(P.n)), dim3(256), 0, s, sa, P.n, (T*)P.pos, (T*)P.vel,
                       (const T*)P.start, (const uint8_t*)P.cand);
    hipLaunchKernelGGL(wrap_escape_kernel<T>, dim3(blocks(P.n)), dim3(256), 0, s, P.g, P.n, (T*)P.pos, P.flag);
    FB_HIP(hipGetLastError());
    if ((rc = scan_ints(s, P.flag, P.rank, P.n, P.sb.bsum, P.sb.total))) return rc;
    hipLaunchKernelGGL(respawn_kernel<T>, dim3(blocks(P.n)), dim3(256), 0, s, P.g, P.n, (const int*)P.flag,
                       (const int*)P.rank, (const int*)P.sb.total, (const Pcg*)P.rng, p.freestream_x,
                       p.freestream_y, (T*)P.pos, (T*)P.vel);
    hipLaunchKernelGGL(rng_advance_kernel, dim3(1), dim3(64), 0, s, P.rng, (const int*)P.sb.total, 0ull);
    FB_HIP(hipGetLastError());
    if ((rc = pic_collide<T>(P, pose, p.wall_omega_rad))) return rc;
  }
  int esc = 0;
  FB_HIP(hipMemcpyAsync(&esc, P.sb.total, sizeof(int), hipMemcpyDeviceToHost, s));
  if ((rc = pic_read_keys(P, 3))) return rc;
  rep->unsupported_faces = (int64_t)P.host_keys[0];
  rep->swept_hits = (int64_t)P.host_keys[2];
  rep->escaped = P.n > 0 ? esc : 0;
  return FB_OK;
}

template <class T>
static int pic_stats(Pic& P, const fb_pose& pose, fb_pic_stats* st) {
  hipStream_t s = P.ctx->stream;
  const int nc = P.g.nx * P.g.ny;
  fill_keys(s, P.keys, 8);
  hipLaunchKernelGGL(fill_int_kernel, dim3(blocks(nc)), dim3(256), 0, s, P.cnt, nc, 0);
  // both kernels fill their slots over the same grid so that one fold covers them
  const unsigned nb = reduction_blocks(std::max((size_t)P.n, (size_t)nc));
  hipLaunchKernelGGL(pic_particle_stats_kernel<T>, dim3(nb), dim3(256), 0, s, P.g, P.n, (const T*)P.pos,
                     (const T*)P.vel, make_naca(P.foil, pose), P.cnt, P.ctx->partials);
  hipLaunchKernelGGL(pic_cell_stats_kernel<T>, dim3(nb), dim3(256), 0, s, nc, (const T*)P.GU, (const T*)P.GV,
                     (const uint8_t*)P.M->solid, (const int*)P.cnt, P.ctx->partials);
  FB_HIP(hipGetLastError());
  FinishSpec f;
  f.add(0, false);  // max particle speed
  f.add(2, true);   // particles inside
  f.add(3, true);   // non-finite (particles)
  f.add(1, false);  // max grid speed
  f.add(3, true);   // non-finite (grid)
  f.add(4, true);   // fluid cells
  f.add(5, true);   // empty cells
  f.add(6, true);   // underfilled cells
  int rc = finish_partials(P.ctx, nb, f, P.keys);
  if (rc) return rc;
  rc = pic_read_keys(P, 7);
  if (rc) return rc;
  auto val = [&](int k) { return P.host_keys[k] ? dkey_inv(P.host_keys[k]) : 0.0; };
  st->max_particle_speed = val(0);
  st->max_grid_speed = val(1);
  st->particles_inside = (int64_t)P.host_keys[2];
  st->nonfinite = (int64_t)P.host_keys[3];
  st->fluid_cells = (int64_t)P.host_keys[4];
  st->empty_cells = (int64_t)P.host_keys[5];
  st->underfilled_cells = (int64_t)P.host_keys[6];
  st->particle_count = P.n;
  return FB_OK;
}

}  // namespace fb

using fb::Pic;

#define PIC_DISPATCH(P, fn, ...) ((P)->prec == FB_F64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

extern "C" {

int fb_pic_create(fb_ctx* ctx, const fb_domain* dom, const fb_foil* foil, int precision, fb_pic** out) {
  FB_REQUIRE(ctx && dom && foil && out, FB_ERR_ARG, "NULL argument");
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  fb::MacGrid* M = nullptr;
  int rc = fb::create_mac(c, dom, foil, precision, &M);
  if (rc) return rc;
  auto* P = new Pic();
  P->ctx = c;
  P->M = M;
  P->g = fb::make_pic_geom(*dom);
  P->foil = *foil;
  P->prec = precision;
  P->sz = precision == FB_F64 ? 8 : 4;
  const size_t nc = (size_t)dom->nx * dom->ny;
  hipError_t e = hipSuccess;
  auto A = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = dev_alloc(p, bytes);
  };
  A(&P->GU, nc * P->sz);
  A(&P->GV, nc * P->sz);
  A(&P->ck_gu, nc * P->sz);
  A(&P->ck_gv, nc * P->sz);
  A(&P->FU, M->NU * P->sz);
  A(&P->FV, M->NV * P->sz);
  A(&P->TU, M->NU * P->sz);
  A(&P->TV, M->NV * P->sz);
  A((void**)&P->flist, nc * 4);
  A((void**)&P->cnt, nc * 4);
  A((void**)&P->rng, sizeof(fb::Pcg));
  A((void**)&P->ck_rng, sizeof(fb::Pcg));
  A((void**)&P->keys, 16 * sizeof(unsigned long long));
  A((void**)&P->table, 2 * fb::kSweepEntries * sizeof(double));
  if (e == hipSuccess) e = host_alloc((void**)&P->host_keys, 16 * sizeof(unsigned long long));
  if (e != hipSuccess) {
    fb::set_error("device allocation failed: %s", hipGetErrorString(e));
    fb::pic_free(P);
    return FB_ERR_NOMEM;
  }
  if ((rc = fb::pic_alloc_particles(*P, 0))) {
    fb::pic_free(P);
    return rc;
  }
  fb::Pcg zero{0ull, 1ull};
  FB_HIP(hipMemcpyAsync(P->rng, &zero, sizeof(zero), hipMemcpyHostToDevice, P->ctx->stream));
  FB_HIP(hipStreamSynchronize(P->ctx->stream));
  *out = reinterpret_cast<fb_pic*>(P);
  return FB_OK;
}

int fb_pic_destroy(fb_pic* pic) {
  if (!pic) return FB_OK;
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  FB_HIP(hipStreamSynchronize(P->ctx->stream));
  fb::pic_free(P);
  return FB_OK;
}

int fb_pic_mac(fb_pic* pic, fb_mac** out) {
  FB_REQUIRE(pic && out, FB_ERR_ARG, "NULL argument");
  *out = reinterpret_cast<fb_mac*>(reinterpret_cast<Pic*>(pic)->M);
  return FB_OK;
}

int fb_pic_set_pose(fb_pic* pic, const fb_pose* pose) {
  FB_REQUIRE(pic && pose, FB_ERR_ARG, "NULL argument");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  return fb::mac_geometry(*P->M, *pose);
}

int fb_pic_set_grid(fb_pic* pic, const void* gu_host, const void* gv_host) {
  FB_REQUIRE(pic && gu_host && gv_host, FB_ERR_ARG, "NULL argument");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  const size_t b = (size_t)P->g.nx * P->g.ny * P->sz;
  FB_HIP(hipMemcpyAsync(P->GU, gu_host, b, hipMemcpyHostToDevice, P->ctx->stream));
  FB_HIP(hipMemcpyAsync(P->GV, gv_host, b, hipMemcpyHostToDevice, P->ctx->stream));
  FB_HIP(hipStreamSynchronize(P->ctx->stream));
  P->grid_valid = true;
  return FB_OK;
}

int fb_pic_get_grid(fb_pic* pic, void* gu_host, void* gv_host) {
  FB_REQUIRE(pic && gu_host && gv_host, FB_ERR_ARG, "NULL argument");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  const size_t b = (size_t)P->g.nx * P->g.ny * P->sz;
  FB_HIP(hipMemcpyAsync(gu_host, P->GU, b, hipMemcpyDeviceToHost, P->ctx->stream));
  FB_HIP(hipMemcpyAsync(gv_host, P->GV, b, hipMemcpyDeviceToHost, P->ctx->stream));
  FB_HIP(hipStreamSynchronize(P->ctx->stream));
  return FB_OK;
}

int fb_pic_set_rng(fb_pic* pic, uint64_t state, uint64_t increment) {
  FB_REQUIRE(pic, FB_ERR_ARG, "NULL handle");
  FB_REQUIRE(increment & 1ull, FB_ERR_ARG, "PCG32 increment must be odd");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  fb::Pcg r{state, increment};
  FB_HIP(hipMemcpyAsync(P->rng, &r, sizeof(r), hipMemcpyHostToDevice, P->ctx->stream));
  FB_HIP(hipStreamSynchronize(P->ctx->stream));
  return FB_OK;
}

int fb_pic_get_rng(fb_pic* pic, uint64_t* state, uint64_t* increment) {
  FB_REQUIRE(pic && state && increment, FB_ERR_ARG, "NULL argument");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  fb::Pcg r{};
  FB_HIP(hipMemcpyAsync(&r, P->rng, sizeof(r), hipMemcpyDeviceToHost, P->ctx->stream));
  FB_HIP(hipStreamSynchronize(P->ctx->stream));
  *state = r.state;
  *increment = r.inc;
  return FB_OK;
}

int fb_pic_seed(fb_pic* pic, int per_cell, int64_t* count) {
  FB_REQUIRE(pic && count, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(per_cell >= 1 && per_cell <= 64, FB_ERR_ARG, "per_cell must be in [1, 64]");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_REQUIRE(P->M->geom_valid, FB_ERR_STATE, "no geometry has been set");
  FB_REQUIRE(P->grid_valid, FB_ERR_STATE, "no grid velocity has been set");
  FB_HIP(hipSetDevice(P->ctx->device));
  return PIC_DISPATCH(P, fb::pic_seed, *P, per_cell, count);
}

int fb_pic_count(fb_pic* pic, int64_t* count) {
  FB_REQUIRE(pic && count, FB_ERR_ARG, "NULL argument");
  *count = reinterpret_cast<Pic*>(pic)->n;
  return FB_OK;
}

int fb_pic_set_particles(fb_pic* pic, int64_t n, const void* pos_host, const void* vel_host) {
  FB_REQUIRE(pic && (n == 0 || (pos_host && vel_host)), FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(n >= 0 && n < (1ll << 30), FB_ERR_ARG, "particle count out of range");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  if (n > P->cap) {
    int rc = fb::pic_alloc_particles(*P, (int)n);
    if (rc) return rc;
  }
  P->n = (int)n;
  const size_t b = (size_t)n * 2 * P->sz;
  if (n) {
    FB_HIP(hipMemcpyAsync(P->pos, pos_host, b, hipMemcpyHostToDevice, P->ctx->stream));
    FB_HIP(hipMemcpyAsync(P->vel, vel_host, b, hipMemcpyHostToDevice, P->ctx->stream));
  }
  FB_HIP(hipStreamSynchronize(P->ctx->stream));
  return FB_OK;
}

int fb_pic_get_particles(fb_pic* pic, void* pos_host, void* vel_host) {
  FB_REQUIRE(pic, FB_ERR_ARG, "NULL handle");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  const size_t b = (size_t)P->n * 2 * P->sz;
  if (P->n && pos_host) FB_HIP(hipMemcpyAsync(pos_host, P->pos, b, hipMemcpyDeviceToHost, P->ctx->stream));
  if (P->n && vel_host) FB_HIP(hipMemcpyAsync(vel_host, P->vel, b, hipMemcpyDeviceToHost, P->ctx->stream));
  FB_HIP(hipStreamSynchronize(P->ctx->stream));
  return FB_OK;
}

int fb_pic_particle_to_grid(fb_pic* pic, int store_grid, double* unsupported_fraction) {
  FB_REQUIRE(pic && unsupported_fraction, FB_ERR_ARG, "NULL argument");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_REQUIRE(P->grid_valid, FB_ERR_STATE, "no grid velocity has been set");
  FB_HIP(hipSetDevice(P->ctx->device));
  fb::fill_keys(P->ctx->stream, P->keys, 1);
  int rc = PIC_DISPATCH(P, fb::pic_p2g, *P, 0);
  if (rc) return rc;
  if (store_grid) {
    rc = P->prec == FB_F64
             ? fb::launch_faces_to_cell<double>(*P->M, (const double*)P->TU, (const double*)P->TV,
                                                (double*)P->GU, (double*)P->GV)
             : fb::launch_faces_to_cell<float>(*P->M, (const float*)P->TU, (const float*)P->TV, (float*)P->GU,
                                               (float*)P->GV);
    if (rc) return rc;
  }
  if ((rc = fb::pic_read_keys(*P, 1))) return rc;
  *unsupported_fraction = (double)P->host_keys[0] / (double)(P->M->NU + P->M->NV);
  return FB_OK;
}

int fb_pic_regrid(fb_pic* pic, const int64_t* indices, int64_t count) {
  FB_REQUIRE(pic && (count == 0 || indices), FB_ERR_ARG, "NULL argument");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_REQUIRE(count >= 0 && count <= P->n, FB_ERR_ARG, "index count out of range");
  if (count == 0) return FB_OK;
  FB_HIP(hipSetDevice(P->ctx->device));
  std::vector<int> idx((size_t)count);
  for (int64_t k = 0; k < count; ++k) {
    FB_REQUIRE(indices[k] >= 0 && indices[k] < P->n, FB_ERR_ARG, "particle index %lld out of range",
               (long long)indices[k]);
    idx[(size_t)k] = (int)indices[k];
  }
  // the flag buffer holds cap + 1 ints and is free between substeps
  FB_HIP(hipMemcpyAsync(P->flag, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, P->ctx->stream));
  hipStream_t s = P->ctx->stream;
  int rc;
  if (P->prec == FB_F64) {
    if ((rc = fb::pic_fallback<double>(*P))) return rc;
    rc = fb::launch_g2p<double>(s, P->g, (int)count, P->flag, (const double*)P->pos, (const double*)P->FU,
                                (const double*)P->FV, (double*)P->vel);
  } else {
    if ((rc = fb::pic_fallback<float>(*P))) return rc;
    rc = fb::launch_g2p<float>(s, P->g, (int)count, P->flag, (const float*)P->pos, (const float*)P->FU,
                               (const float*)P->FV, (float*)P->vel);
  }
  if (rc) return rc;
  FB_HIP(hipStreamSynchronize(s));
  return FB_OK;
}

int fb_pic_resolve_collisions(fb_pic* pic, const fb_pose* pose, double wall_omega_rad) {
  FB_REQUIRE(pic && pose, FB_ERR_ARG, "NULL argument");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  int rc = PIC_DISPATCH(P, fb::pic_collide, *P, *pose, wall_omega_rad);
  if (rc) return rc;
  FB_HIP(hipStreamSynchronize(P->ctx->stream));
  return FB_OK;
}

int fb_pic_substep(fb_pic* pic, const fb_pic_params* params, const fb_pose* pose, const double* sweep_table,
                   fb_pic_report* report) {
  FB_REQUIRE(pic && params && pose && sweep_table && report, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(params->max_iterations >= 1, FB_ERR_ARG, "max_iterations must be positive");
  FB_REQUIRE(params->mixed_poisson >= 0 && params->mixed_poisson <= 2, FB_ERR_ARG, "unknown pressure precision %d",
             (int)params->mixed_poisson);
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_REQUIRE(P->grid_valid, FB_ERR_STATE, "no grid velocity has been set");
  FB_HIP(hipSetDevice(P->ctx->device));
  *report = fb_pic_report{};
  return PIC_DISPATCH(P, fb::pic_substep, *P, *params, *pose, sweep_table, report);
}

int fb_pic_get_stats(fb_pic* pic, const fb_pose* pose, fb_pic_stats* stats) {
  FB_REQUIRE(pic && pose && stats, FB_ERR_ARG, "NULL argument");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_REQUIRE(P->M->geom_valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(P->ctx->device));
  return PIC_DISPATCH(P, fb::pic_stats, *P, *pose, stats);
}

int fb_pic_checkpoint(fb_pic* pic) {
  FB_REQUIRE(pic, FB_ERR_ARG, "NULL handle");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  hipStream_t s = P->ctx->stream;
  const size_t pb = (size_t)P->n * 2 * P->sz, gb = (size_t)P->g.nx * P->g.ny * P->sz;
  if (pb) {
    FB_HIP(hipMemcpyAsync(P->ck_pos, P->pos, pb, hipMemcpyDeviceToDevice, s));
    FB_HIP(hipMemcpyAsync(P->ck_vel, P->vel, pb, hipMemcpyDeviceToDevice, s));
  }
  FB_HIP(hipMemcpyAsync(P->ck_gu, P->GU, gb, hipMemcpyDeviceToDevice, s));
  FB_HIP(hipMemcpyAsync(P->ck_gv, P->GV, gb, hipMemcpyDeviceToDevice, s));
  FB_HIP(hipMemcpyAsync(P->ck_rng, P->rng, sizeof(fb::Pcg), hipMemcpyDeviceToDevice, s));
  P->ck_n = P->n;
  int rc = fb_mac_checkpoint(reinterpret_cast<fb_mac*>(P->M));
  if (rc) return rc;
  P->ck_valid = true;
  return FB_OK;
}

int fb_pic_rollback(fb_pic* pic) {
  FB_REQUIRE(pic, FB_ERR_ARG, "NULL handle");
  auto* P = reinterpret_cast<Pic*>(pic);
  FB_REQUIRE(P->ck_valid, FB_ERR_STATE, "no checkpoint to roll back to");
  FB_HIP(hipSetDevice(P->ctx->device));
  hipStream_t s = P->ctx->stream;
  P->n = P->ck_n;
  const size_t pb = (size_t)P->n * 2 * P->sz, gb = (size_t)P->g.nx * P->g.ny * P->sz;
  if (pb) {
    FB_HIP(hipMemcpyAsync(P->pos, P->ck_pos, pb, hipMemcpyDeviceToDevice, s));
    FB_HIP(hipMemcpyAsync(P->vel, P->ck_vel, pb, hipMemcpyDeviceToDevice, s));
  }
  FB_HIP(hipMemcpyAsync(P->GU, P->ck_gu, gb, hipMemcpyDeviceToDevice, s));
  FB_HIP(hipMemcpyAsync(P->GV, P->ck_gv, gb, hipMemcpyDeviceToDevice, s));
  FB_HIP(hipMemcpyAsync(P->rng, P->ck_rng, sizeof(fb::Pcg), hipMemcpyDeviceToDevice, s));
  return fb_mac_rollback(reinterpret_cast<fb_mac*>(P->M));
}

}  // extern "C"

// ------------------------------------------------------------------ standalone transfers

namespace fb {

template <class T>
static int standalone_p2f(Ctx* c, int64_t n, const void* pos, const void* vel, const void* fu, const void* fv,
                          const PicGeom& g, void* uo, void* vo, double* frac) {
  hipStream_t s = c->stream;
  const size_t nu = (size_t)g.ny * (g.nx + 1), nv = (size_t)(g.ny + 1) * g.nx;
  SortBufs b;
  int rc = alloc_sort(b, g.nx * g.ny, (int)std::max<int64_t>(n, 1), sizeof(T), s, true);
  if (rc) return rc;
  void *dp = nullptr, *dv = nullptr, *dfu = nullptr, *dfv = nullptr, *du = nullptr, *dvv = nullptr;
  unsigned long long* key = nullptr;
  const size_t pb = (size_t)std::max<int64_t>(n, 1) * 2 * sizeof(T);
  hipError_t e = hipSuccess;
  for (auto pr : {std::make_pair(&dp, pb), std::make_pair(&dv, pb), std::make_pair(&dfu, nu * sizeof(T)),
                  std::make_pair(&dfv, nv * sizeof(T)), std::make_pair(&du, nu * sizeof(T)),
                  std::make_pair(&dvv, nv * sizeof(T))})
    if (e == hipSuccess) e = hipMallocAsync(pr.first, pr.second, s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&key, 8, s);
  if (e == hipSuccess && n) e = hipMemcpyAsync(dp, pos, (size_t)n * 2 * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && n) e = hipMemcpyAsync(dv, vel, (size_t)n * 2 * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dfu, fu, nu * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dfv, fv, nv * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    fill_keys(s, key, 1);
    rc = sort_by_cell<T>(s, g, (int)n, (const T*)dp, (const T*)dv, b);
    if (!rc) rc = launch_p2g<T>(s, g, b, (const T*)dfu, (const T*)dfv, (T*)du, (T*)dvv, key);
    unsigned long long uns = 0;
    if (!rc) {
      e = hipMemcpyAsync(uo, du, nu * sizeof(T), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(vo, dvv, nv * sizeof(T), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(&uns, key, 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      *frac = (double)uns / (double)(nu + nv);
    }
  }
  for (void* p : {dp, dv, dfu, dfv, du, dvv, (void*)key})
    if (p) (void)hipFreeAsync(p, s);
  free_sort(b, s, true);
  if (e != hipSuccess) {
    set_error("particle_to_faces failed: %s", hipGetErrorString(e));
    return FB_ERR_HIP;
  }
  FB_HIP(hipStreamSynchronize(s));
  return rc;
}

template <class T>
static int standalone_f2p(Ctx* c, int64_t n, const void* u, const void* v, const void* pos, const PicGeom& g,
                          void* out) {
  hipStream_t s = c->stream;
  const size_t nu = (size_t)g.ny * (g.nx + 1), nv = (size_t)(g.ny + 1) * g.nx;
  void *dp = nullptr, *du = nullptr, *dv = nullptr, *dout = nullptr;
  const size_t pb = (size_t)std::max<int64_t>(n, 1) * 2 * sizeof(T);
  hipError_t e = hipMallocAsync(&dp, pb, s);
  if (e == hipSuccess) e = hipMallocAsync(&dout, pb, s);
  if (e == hipSuccess) e = hipMallocAsync(&du, nu * sizeof(T), s);
  if (e == hipSuccess) e = hipMallocAsync(&dv, nv * sizeof(T), s);
  if (e == hipSuccess && n) e = hipMemcpyAsync(dp, pos, (size_t)n * 2 * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(du, u, nu * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dv, v, nv * sizeof(T), hipMemcpyHostToDevice, s);
  int rc = FB_OK;
  if (e == hipSuccess) {
    rc = launch_g2p<T>(s, g, (int)n, nullptr, (const T*)dp, (const T*)du, (const T*)dv, (T*)dout);
    if (!rc && n) e = hipMemcpyAsync(out, dout, (size_t)n * 2 * sizeof(T), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  for (void* p : {dp, du, dv, dout})
    if (p) (void)hipFreeAsync(p, s);
  if (e != hipSuccess) {
    set_error("faces_to_particle failed: %s", hipGetErrorString(e));
    return FB_ERR_HIP;
  }
  FB_HIP(hipStreamSynchronize(s));
  return rc;
}

template <class T>
static int standalone_p2c(Ctx* c, int64_t n, const void* pos, const void* vel, const PicGeom& g, double fsx,
                          double fsy, void* out) {
  hipStream_t s = c->stream;
  const size_t gb = (size_t)g.nx * g.ny * 2 * sizeof(T);
  const int nbin = base_bins(g.nx, g.px) * base_bins(g.ny, g.py);
  SortBufs b;
  int rc = alloc_sort(b, nbin, (int)std::max<int64_t>(n, 1), sizeof(T), s, true);
  if (rc) return rc;
  void *dp = nullptr, *dv = nullptr, *dg = nullptr;
  const size_t pb = (size_t)std::max<int64_t>(n, 1) * 2 * sizeof(T);
  hipError_t e = hipMallocAsync(&dp, pb, s);
  if (e == hipSuccess) e = hipMallocAsync(&dv, pb, s);
  if (e == hipSuccess) e = hipMallocAsync(&dg, gb, s);
  if (e == hipSuccess && n) e = hipMemcpyAsync(dp, pos, (size_t)n * 2 * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && n) e = hipMemcpyAsync(dv, vel, (size_t)n * 2 * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    rc = sort_by_cell<T, true>(s, g, (int)n, (const T*)dp, (const T*)dv, b);
    if (!rc) {
      hipLaunchKernelGGL(p2c_kernel<T>, dim3((g.nx + 255) / 256, g.ny), dim3(256), 0, s, g, b.cstart, b.sidx,
                         (const T*)b.spos, (const T*)b.sval, fsx, fsy, (T*)dg);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(out, dg, gb, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
  }
  for (void* p : {dp, dv, dg})
    if (p) (void)hipFreeAsync(p, s);
  free_sort(b, s, true);
  if (e != hipSuccess) {
    set_error("particle_to_cells failed: %s", hipGetErrorString(e));
    return FB_ERR_HIP;
  }
  FB_HIP(hipStreamSynchronize(s));
  return rc;
}

template <class T>
static int standalone_c2p(Ctx* c, int64_t n, const void* grid, const void* pos, const PicGeom& g, void* out) {
  hipStream_t s = c->stream;
  const size_t gb = (size_t)g.nx * g.ny * 2 * sizeof(T);
  void *dp = nullptr, *dg = nullptr, *dout = nullptr;
  const size_t pb = (size_t)std::max<int64_t>(n, 1) * 2 * sizeof(T);
  hipError_t e = hipMallocAsync(&dp, pb, s);
  if (e == hipSuccess) e = hipMallocAsync(&dout, pb, s);
  if (e == hipSuccess) e = hipMallocAsync(&dg, gb, s);
  if (e == hipSuccess && n) e = hipMemcpyAsync(dp, pos, (size_t)n * 2 * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dg, grid, gb, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && n > 0) {
    hipLaunchKernelGGL(c2p_kernel<T>, dim3(blocks(n)), dim3(256), 0, s, g, (int)n, (const T*)dg, (const T*)dp,
                       (T*)dout);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, (size_t)n * 2 * sizeof(T), hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (void* p : {dp, dg, dout})
    if (p) (void)hipFreeAsync(p, s);
  if (e != hipSuccess) {
    set_error("cells_to_particle failed: %s", hipGetErrorString(e));
    return FB_ERR_HIP;
  }
  FB_HIP(hipStreamSynchronize(s));
  return FB_OK;
}

// ------------------------------------------------------------------ allocation-free P2G
// fb_pic_particle_to_faces_noalloc keeps its counting-sort buffers (cell of each particle, per-cell counts,
// starts and cursors, the scan's block sums, the sorted index and the sorted positions and values) in one
// file-scope device array of fixed size, laid out for the largest particle and cell counts it accepts. Each
// GPU holds its own copy until the process ends. Host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the array's name is a host-side stand-in).
constexpr int64_t kNoAllocMaxParticles = 1 << 19;  // 524,288 particles
constexpr int64_t kNoAllocMaxCells = 1 << 18;      // 262,144 cells (512 x 512)
constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNoAllocScanBlocks =
    (size_t)((kNoAllocMaxCells > kNoAllocMaxParticles ? kNoAllocMaxCells : kNoAllocMaxParticles) + 1023) / 1024 + 2;
constexpr size_t kNoAllocBytes = 2 * al256(kNoAllocMaxParticles * 4)         // cellof, sidx
                                 + 2 * al256((kNoAllocMaxCells + 1) * 4)      // counts, cstart
                                 + al256(kNoAllocMaxCells * 4)                // cursor
                                 + al256(kNoAllocScanBlocks * 4) + al256(16)  // bsum, total
                                 + 2 * al256(kNoAllocMaxParticles * 2 * 8);   // spos, sval (FP64 worst case)

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kNoAllocMaxDevices] = {};

// Device address of `scratch` on `device` (current on the calling thread), looked up on the first call per
// device and kept.
static int scratch_address(int device, void** out) {
  FB_REQUIRE(device >= 0 && device < kNoAllocMaxDevices, FB_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
      return FB_ERR_HIP;
    }
    g_scratch_dev[device] = p;
  }
  *out = g_scratch_dev[device];
  return FB_OK;
}

// The sort buffers carved out of the scratch array at fixed offsets.
static SortBufs scratch_sort(void* base, int ncell, int cap) {
  unsigned char* p = static_cast<unsigned char*>(base);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    void* r = p + off;
    off += al256(bytes);
    return r;
  };
  SortBufs b;
  b.ncell = ncell;
  b.cap = cap;
  b.cellof = (int*)take(kNoAllocMaxParticles * 4);
  b.sidx = (int*)take(kNoAllocMaxParticles * 4);
  b.counts = (int*)take((kNoAllocMaxCells + 1) * 4);
  b.cstart = (int*)take((kNoAllocMaxCells + 1) * 4);
  b.cursor = (int*)take(kNoAllocMaxCells * 4);
  b.bsum = (int*)take(kNoAllocScanBlocks * 4);
  b.total = (int*)take(16);
  b.spos = take(kNoAllocMaxParticles * 2 * 8);
  b.sval = take(kNoAllocMaxParticles * 2 * 8);
  return b;
}

template <class T>
static int noalloc_p2f(Ctx* c, int64_t n, const void* pos, const void* vel, const void* fu, const void* fv,
                       const PicGeom& g, void* uo, void* vo, uint64_t* unsupported) {
  void* base = nullptr;
  int rc = scratch_address(c->device, &base);
  if (rc) return rc;
  hipStream_t s = c->stream;
  SortBufs b = scratch_sort(base, g.nx * g.ny, (int)std::max<int64_t>(n, 1));
  // the scratch keeps the previous call's contents: the counts and cursors are cleared by sort_by_cell's
  // fill kernels, the unsupported-face count here (kernels, not memsets, so the calls can be captured)
  fill_keys(s, reinterpret_cast<unsigned long long*>(unsupported), 1);
  rc = sort_by_cell<T>(s, g, (int)n, (const T*)pos, (const T*)vel, b);
  if (!rc)
    rc = launch_p2g<T>(s, g, b, (const T*)fu, (const T*)fv, (T*)uo, (T*)vo,
                       reinterpret_cast<unsigned long long*>(unsupported));
  return rc;
}

static PicGeom transfer_geom(int nx, int ny, double x0, double y0, double dx, double dy, int px, int py) {
  PicGeom g;
  g.nx = nx;
  g.ny = ny;
  g.px = px;
  g.py = py;
  g.x0 = x0;
  g.y0 = y0;
  g.dx = dx;
  g.dy = dy;
  g.x1 = x0 + nx * dx;
  g.y1 = y0 + ny * dy;
  return g;
}

}  // namespace fb

extern "C" {

int fb_pic_particle_to_faces(fb_ctx* ctx, int precision, int64_t n, const void* positions, const void* velocities,
                             const void* fallback_u, const void* fallback_v, int nx, int ny, double x0, double y0,
                             double dx, double dy, int periodic_x, int periodic_y, void* u_out, void* v_out,
                             double* unsupported_fraction) {
  FB_REQUIRE(ctx && fallback_u && fallback_v && u_out && v_out && unsupported_fraction, FB_ERR_ARG,
             "NULL argument");
  FB_REQUIRE(n == 0 || (positions && velocities), FB_ERR_ARG, "NULL particle arrays");
  FB_REQUIRE(n >= 0 && n < (1ll << 30), FB_ERR_ARG, "particle count out of range");
  FB_REQUIRE(nx >= 4 && ny >= 4, FB_ERR_ARG, "each face grid axis must span at least four cells");
  FB_REQUIRE(precision == FB_F64 || precision == FB_F32, FB_ERR_ARG, "precision must be FB_F64 or FB_F32");
  FB_REQUIRE(dx > 0.0 && dy > 0.0, FB_ERR_ARG, "cell sizes must be positive");
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  fb::PicGeom g = fb::transfer_geom(nx, ny, x0, y0, dx, dy, periodic_x != 0, periodic_y != 0);
  return precision == FB_F64
             ? fb::standalone_p2f<double>(c, n, positions, velocities, fallback_u, fallback_v, g, u_out, v_out,
                                          unsupported_fraction)
             : fb::standalone_p2f<float>(c, n, positions, velocities, fallback_u, fallback_v, g, u_out, v_out,
                                         unsupported_fraction);
}

int fb_pic_particle_to_faces_noalloc(fb_ctx* ctx, int precision, int64_t n, const void* positions,
                                     const void* velocities, const void* fallback_u, const void* fallback_v, int nx,
                                     int ny, double x0, double y0, double dx, double dy, int periodic_x,
                                     int periodic_y, void* u_out, void* v_out, uint64_t* unsupported_faces) {
  FB_REQUIRE(ctx && fallback_u && fallback_v && u_out && v_out && unsupported_faces, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(n == 0 || (positions && velocities), FB_ERR_ARG, "NULL particle arrays");
  FB_REQUIRE(n >= 0, FB_ERR_ARG, "negative particle count");
  FB_REQUIRE(n <= fb::kNoAllocMaxParticles, FB_ERR_ARG,
             "%lld particles exceed the allocation-free maximum of %lld", (long long)n,
             (long long)fb::kNoAllocMaxParticles);
  FB_REQUIRE(nx >= 4 && ny >= 4, FB_ERR_ARG, "each face grid axis must span at least four cells");
  FB_REQUIRE((int64_t)nx * ny <= fb::kNoAllocMaxCells, FB_ERR_ARG,
             "%d x %d cells exceed the allocation-free maximum of %lld", nx, ny, (long long)fb::kNoAllocMaxCells);
  FB_REQUIRE(precision == FB_F64 || precision == FB_F32, FB_ERR_ARG, "precision must be FB_F64 or FB_F32");
  FB_REQUIRE(dx > 0.0 && dy > 0.0, FB_ERR_ARG, "cell sizes must be positive");
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  fb::PicGeom g = fb::transfer_geom(nx, ny, x0, y0, dx, dy, periodic_x != 0, periodic_y != 0);
  return precision == FB_F64
             ? fb::noalloc_p2f<double>(c, n, positions, velocities, fallback_u, fallback_v, g, u_out, v_out,
                                       unsupported_faces)
             : fb::noalloc_p2f<float>(c, n, positions, velocities, fallback_u, fallback_v, g, u_out, v_out,
                                      unsupported_faces);
}

int fb_pic_faces_to_particle(fb_ctx* ctx, int precision, int64_t n, const void* u, const void* v,
                             const void* positions, int nx, int ny, double x0, double y0, double dx, double dy,
                             int periodic_x, int periodic_y, void* out) {
  FB_REQUIRE(ctx && u && v, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(n == 0 || (positions && out), FB_ERR_ARG, "NULL particle arrays");
  FB_REQUIRE(n >= 0 && n < (1ll << 30), FB_ERR_ARG, "particle count out of range");
  FB_REQUIRE(nx >= 1 && ny >= 1, FB_ERR_ARG, "empty face grid");
  FB_REQUIRE(precision == FB_F64 || precision == FB_F32, FB_ERR_ARG, "precision must be FB_F64 or FB_F32");
  FB_REQUIRE(dx > 0.0 && dy > 0.0, FB_ERR_ARG, "cell sizes must be positive");
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  fb::PicGeom g = fb::transfer_geom(nx, ny, x0, y0, dx, dy, periodic_x != 0, periodic_y != 0);
  return precision == FB_F64 ? fb::standalone_f2p<double>(c, n, u, v, positions, g, out)
                             : fb::standalone_f2p<float>(c, n, u, v, positions, g, out);
}

int fb_pic_particle_to_cells(fb_ctx* ctx, int precision, int64_t n, const void* positions, const void* velocities,
                             int nx, int ny, double x0, double y0, double dx, double dy, double freestream_x,
                             double freestream_y, int periodic_x, int periodic_y, void* grid_out) {
  FB_REQUIRE(ctx && grid_out, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(n == 0 || (positions && velocities), FB_ERR_ARG, "NULL particle arrays");
  FB_REQUIRE(n >= 0 && n < (1ll << 30), FB_ERR_ARG, "particle count out of range");
  FB_REQUIRE(nx > 0 && ny > 0 && dx > 0.0 && dy > 0.0, FB_ERR_ARG, "grid spacing and resolution must be positive");
  FB_REQUIRE(ny < 65536 && (int64_t)nx * ny <= (1ll << 28), FB_ERR_ARG, "grid of %d x %d cells is too large", nx,
             ny);
  FB_REQUIRE(precision == FB_F64 || precision == FB_F32, FB_ERR_ARG, "precision must be FB_F64 or FB_F32");
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  fb::PicGeom g = fb::transfer_geom(nx, ny, x0, y0, dx, dy, periodic_x != 0, periodic_y != 0);
  return precision == FB_F64
             ? fb::standalone_p2c<double>(c, n, positions, velocities, g, freestream_x, freestream_y, grid_out)
             : fb::standalone_p2c<float>(c, n, positions, velocities, g, freestream_x, freestream_y, grid_out);
}

int fb_pic_cells_to_particle(fb_ctx* ctx, int precision, int64_t n, const void* grid, const void* positions, int nx,
                             int ny, double x0, double y0, double dx, double dy, int periodic_x, int periodic_y,
                             void* velocities_out) {
  FB_REQUIRE(ctx && grid, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(n == 0 || (positions && velocities_out), FB_ERR_ARG, "NULL particle arrays");
  FB_REQUIRE(n >= 0 && n < (1ll << 30), FB_ERR_ARG, "particle count out of range");
  FB_REQUIRE(nx > 0 && ny > 0, FB_ERR_ARG, "grid axes must be nonempty");
  FB_REQUIRE((int64_t)nx * ny <= (1ll << 28), FB_ERR_ARG, "grid of %d x %d cells is too large", nx, ny);
  FB_REQUIRE(dx > 0.0 && dy > 0.0, FB_ERR_ARG, "grid spacing must be positive");
  FB_REQUIRE(precision == FB_F64 || precision == FB_F32, FB_ERR_ARG, "precision must be FB_F64 or FB_F32");
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  fb::PicGeom g = fb::transfer_geom(nx, ny, x0, y0, dx, dy, periodic_x != 0, periodic_y != 0);
  return precision == FB_F64 ? fb::standalone_c2p<double>(c, n, grid, positions, g, velocities_out)
                             : fb::standalone_c2p<float>(c, n, grid, positions, g, velocities_out);
}

// The exported grid velocity of the handle (the cell planes of the last grid update) and its mask.
static fb::DiagField pic_diag_field(const fb::Pic& P) {
  fb::DiagField f;
  f.nx = P.g.nx;
  f.ny = P.g.ny;
  f.px = P.g.px;
  f.py = P.g.py;
  f.x0 = P.g.x0;
  f.y0 = P.g.y0;
  f.dx = P.g.dx;
  f.dy = P.g.dy;
  f.kind = fb::kDiagCells;
  f.storage_f64 = f.arith_f64 = P.sz == sizeof(double);
  f.u = P.GU;
  f.v = P.GV;
  f.stride = 1;
  f.solid = P.M->solid;
  return f;
}

int fb_pic_diagnostics(fb_pic* pic, const fb_diag_params* params, fb_diag* out, int64_t* histogram,
                       int32_t* counts) {
  FB_REQUIRE(pic && params && out && histogram, FB_ERR_ARG, "NULL argument");
  auto* P = reinterpret_cast<fb::Pic*>(pic);
  FB_REQUIRE(P->M->geom_valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(P->ctx->device));
  *out = fb_diag{};
  const fb::DiagField f = pic_diag_field(*P);
  int rc = fb::field_diagnostics(P->ctx, f, *params, out);
  if (!rc) rc = fb::particle_census(P->ctx, f, P->n, P->pos, out, histogram, counts);
  return rc;
}

int fb_pic_sample_velocity(fb_pic* pic, int64_t n, const double* points, void* out) {
  FB_REQUIRE(pic, FB_ERR_ARG, "NULL handle");
  auto* P = reinterpret_cast<fb::Pic*>(pic);
  FB_HIP(hipSetDevice(P->ctx->device));
  return fb::sample_cell_velocity(P->ctx, pic_diag_field(*P), n, points, out);
}

}  // extern "C"
