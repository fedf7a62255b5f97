// Staggered MAC grid on gfx950: the Stable Fluids substep (fb/solvers/stable_fluids.py:424-494)
// and the grid pieces PIC/FLIP shares (fb/core/grid.py).
//
// Layout: faces u (ny, nx+1) and v (ny+1, nx) row-major in the storage dtype; cell fields as
// separate x/y planes (ny, nx). The pressure solve is FP64 (or FP32 sweeps with FP64
// refinement in the mixed form) on (ny, nx). Arithmetic follows the reference's NumPy dtype
// rules: Python-float scalars are rounded to the storage dtype, bilinear weights are FP64
// (fb/core/interpolation.py:9-52), and the pressure gradient is applied in FP64 then rounded.
#include <algorithm>
#include <vector>

#include "diag.h"
#include "mac.h"

namespace fb {

#pragma clang fp contract(off)

__device__ __forceinline__ double cx_at(const MacGeom& g, int i) { return linspace_at(i, g.nx, g.sx, g.stx, g.ex); }
__device__ __forceinline__ double cy_at(const MacGeom& g, int j) { return linspace_at(j, g.ny, g.sy, g.sty, g.ey); }

// np.mod for floats (Python sign convention).
__device__ __forceinline__ double pymod(double a, double b) {
  double m = fmod(a, b);
  if (m != 0.0) {
    if ((b < 0.0) != (m < 0.0)) m += b;
  } else {
    m = copysign(0.0, b);
  }
  return m;
}

// sample_staggered_scalar with offset (0.5, 0.5) on a (ny, nx) cell field
// (fb/core/interpolation.py:9-52): FP64 weights, result rounded to T.
template <class T>
__device__ __forceinline__ T sample_cell(const MacGeom& g, const T* __restrict__ f, double px, double py) {
  double gx = (px - g.x0) / g.dx - 0.5;
  double gy = (py - g.y0) / g.dy - 0.5;
  gx = g.px ? pymod(gx, (double)g.nx) : fmin(fmax(gx, 0.0), (double)g.nx - 1.0);
  gy = g.py ? pymod(gy, (double)g.ny) : fmin(fmax(gy, 0.0), (double)g.ny - 1.0);
  double fx = floor(gx), fy = floor(gy);
  int ix0 = (int)fx, iy0 = (int)fy;
  double tx = gx - fx, ty = gy - fy;
  // in-bounds guard: np.mod can round up to the period for tiny negative arguments
  ix0 = min(max(ix0, 0), g.nx - 1);
  iy0 = min(max(iy0, 0), g.ny - 1);
  int ix1 = g.px ? (ix0 + 1) % g.nx : min(ix0 + 1, g.nx - 1);
  int iy1 = g.py ? (iy0 + 1) % g.ny : min(iy0 + 1, g.ny - 1);
  double a = (double)f[(size_t)iy0 * g.nx + ix0], b = (double)f[(size_t)iy0 * g.nx + ix1];
  double c = (double)f[(size_t)iy1 * g.nx + ix0], d = (double)f[(size_t)iy1 * g.nx + ix1];
  double s = (1.0 - tx) * (1.0 - ty) * a + tx * (1.0 - ty) * b + (1.0 - tx) * ty * c + tx * ty * d;
  return (T)s;
}

// rk2_backtrace (fb/core/grid.py:527-537) from point (px, py) through the field (fu, fv).
template <class T>
__device__ __forceinline__ void backtrace(const MacGeom& g, const T* fu, const T* fv, double px, double py, double dt,
                                          double& ox, double& oy) {
  T u0 = sample_cell(g, fu, px, py), v0 = sample_cell(g, fv, px, py);
  T h = (T)(0.5 * dt);
  double mx = px - (double)(h * u0);
  double my = py - (double)(h * v0);
  T um = sample_cell(g, fu, mx, my), vm = sample_cell(g, fv, mx, my);
  T d = (T)dt;
  ox = px - (double)(d * um);
  oy = py - (double)(d * vm);
}

// faces_to_cell (fb/core/grid.py:25-29)
template <class T>
__global__ void __launch_bounds__(256) faces_to_cell_kernel(int nx, int ny, const T* __restrict__ u,
                                                            const T* __restrict__ v, T* __restrict__ cu,
                                                            T* __restrict__ cv) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= nx) return;
  size_t i = (size_t)y * nx + x;
  cu[i] = (T)0.5 * (u[(size_t)y * (nx + 1) + x] + u[(size_t)y * (nx + 1) + x + 1]);
  cv[i] = (T)0.5 * (v[(size_t)y * nx + x] + v[(size_t)(y + 1) * nx + x]);
}

// cell_to_faces (fb/core/grid.py:32-44)
template <class T>
__global__ void __launch_bounds__(256) cell_to_faces_kernel(int nx, int ny, const T* __restrict__ cu,
                                                            const T* __restrict__ cv, T* __restrict__ u,
                                                            T* __restrict__ v) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (y < ny && x <= nx) {
    T val;
    if (x == 0) val = cu[(size_t)y * nx];
    else if (x == nx) val = cu[(size_t)y * nx + nx - 1];
    else val = (T)0.5 * (cu[(size_t)y * nx + x - 1] + cu[(size_t)y * nx + x]);
    u[(size_t)y * (nx + 1) + x] = val;
  }
  if (y <= ny && x < nx) {
    T val;
    if (y == 0) val = cv[x];
    else if (y == ny) val = cv[(size_t)(ny - 1) * nx + x];
    else val = (T)0.5 * (cv[(size_t)(y - 1) * nx + x] + cv[(size_t)y * nx + x]);
    v[(size_t)y * nx + x] = val;
  }
}

// First semi-Lagrangian pass (fb/core/grid.py:512-515).
template <class T>
__global__ void __launch_bounds__(256) advect_first_kernel(MacGeom g, const T* __restrict__ cu,
                                                           const T* __restrict__ cv, double dt, T* __restrict__ fu,
                                                           T* __restrict__ fv) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= g.nx) return;
  double px = cx_at(g, x), py = cy_at(g, y), dxp, dyp;
  backtrace(g, cu, cv, px, py, dt, dxp, dyp);
  size_t i = (size_t)y * g.nx + x;
  fu[i] = sample_cell(g, cu, dxp, dyp);
  fv[i] = sample_cell(g, cv, dxp, dyp);
}

// Limited MacCormack correction (fb/core/grid.py:518-524, 540-569).
template <class T>
__global__ void __launch_bounds__(256) advect_correct_kernel(MacGeom g, const T* __restrict__ cu,
                                                             const T* __restrict__ cv, const T* __restrict__ fu,
                                                             const T* __restrict__ fv, double dt, T* __restrict__ ou,
                                                             T* __restrict__ ov) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= g.nx) return;
  double px = cx_at(g, x), py = cy_at(g, y), bx, by;
  backtrace(g, fu, fv, px, py, -dt, bx, by);
  T wu = sample_cell(g, fu, bx, by), wv = sample_cell(g, fv, bx, by);
  size_t i = (size_t)y * g.nx + x;
  T fu0 = fu[i], fv0 = fv[i];
  T cu0 = cu[i], cv0 = cv[i];
  T corr_u = fu0 + (T)0.5 * (cu0 - wu);
  T corr_v = fv0 + (T)0.5 * (cv0 - wv);
  T lo_u = cu0, hi_u = cu0, lo_v = cv0, hi_v = cv0;
  for (int oy = -1; oy <= 1; ++oy) {
    int r = g.py ? (y + oy + g.ny) % g.ny : min(max(y + oy, 0), g.ny - 1);
    for (int ox = -1; ox <= 1; ++ox) {
      int c = g.px ? (x + ox + g.nx) % g.nx : min(max(x + ox, 0), g.nx - 1);
      size_t k = (size_t)r * g.nx + c;
      T a = cu[k], b = cv[k];
      lo_u = a < lo_u ? a : lo_u;
      hi_u = a > hi_u ? a : hi_u;
      lo_v = b < lo_v ? b : lo_v;
      hi_v = b > hi_v ? b : hi_v;
    }
  }
  lo_u = fu0 < lo_u ? fu0 : lo_u;
  hi_u = fu0 > hi_u ? fu0 : hi_u;
  lo_v = fv0 < lo_v ? fv0 : lo_v;
  hi_v = fv0 > hi_v ? fv0 : hi_v;
  T ru = corr_u < lo_u ? lo_u : corr_u;
  ru = ru > hi_u ? hi_u : ru;
  T rv = corr_v < lo_v ? lo_v : corr_v;
  rv = rv > hi_v ? hi_v : rv;
  ou[i] = ru;
  ov[i] = rv;
}

// Max |a| over a field into keys[slot] (order-free, exact).
// Per-block max |a| into partial slot q (grid-stride over n values).
template <class T>
__global__ void __launch_bounds__(256) absmax_kernel(size_t n, const T* __restrict__ a,
                                                     unsigned long long* __restrict__ part, int q) {
  __shared__ unsigned long long lds[4];
  unsigned long long m = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const unsigned long long k = mag_key((double)a[i]);
    m = k > m ? k : m;
  }
  put_max(part, q, m, lds);
}

// One Jacobi sweep of (I - nu dt Lap) u = u0 (fb/core/grid.py:245-267) on a logical W x H
// field (row pitch P), edge replication or periodic wrap; blockIdx.z selects the component.
// Sweep k (1-based) is skipped once the component's sweep k-1 met the tolerance.
template <class T>
struct JacobiArgs {
  const T* in[2];
  const T* orig[2];
  T* out[2];
  int W[2], H[2], P[2];
  int px, py;
  double ax, ay, den;
  const unsigned long long* scale_keys;  // keys of max |original| per component
  double tol;
  int k;
  unsigned long long* resid;  // [comp][maxit+1] keys of max |upd - res|
  int direct;                 // blocks fold into resid themselves (else per-block partials and a fold launch)
  int stride;
  unsigned long long* part;  // per-block maxima of this sweep, slot = component
};

// max(max |u0|, 1) of fb/core/grid.py:255 (NaN propagates).
__device__ __forceinline__ double jacobi_scale(unsigned long long key) {
  const double m = key ? dkey_inv(key) : 0.0;
  return (m != m) ? m : fmax(m, 1.0);
}

template <class T>
__global__ void __launch_bounds__(256) jacobi_sweep_kernel(JacobiArgs<T> a) {
  __shared__ unsigned long long lds[4];
  const int c = blockIdx.z;
  if (a.k > 1) {
    const unsigned long long key = a.resid[c * a.stride + a.k - 1];
    if (key != 0ull && dkey_inv(key) / jacobi_scale(a.scale_keys[c]) <= a.tol) {
      // converged: carry the residual forward so every later sweep of this solve skips too
      // (a later sweep would otherwise read a ping-pong buffer this solve never wrote)
      if (a.direct) {
        if (threadIdx.x == 0 && blockIdx.x == 0) key_max(&a.resid[c * a.stride + a.k], key);
      } else if (threadIdx.x == 0) {
        a.part[(size_t)c * kPartialBlocks + blockIdx.x] = key;
      }
      return;
    }
  }
  const int W = a.W[c], H = a.H[c], P = a.P[c];
  const T* __restrict__ r = a.in[c];
  const T* __restrict__ orig = a.orig[c];
  T* __restrict__ out = a.out[c];
  const T ax = (T)a.ax, ay = (T)a.ay, den = (T)a.den;
  unsigned long long m = 0;
  const size_t n = (size_t)W * H;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (size_t)gridDim.x * 256) {
    // 32-bit index arithmetic (a grid has far fewer than 2^32 cells; a 64-bit division is a long
    // software sequence on the GPU)
    const unsigned ui = (unsigned)idx;
    const int y = (int)(ui / (unsigned)W), x = (int)(ui - (unsigned)y * (unsigned)W);
    auto at = [&](int xx, int yy) { return r[(size_t)yy * P + xx]; };
    // neighbours: wrapped when periodic, else the edge replicated (selects rather than a modulo)
    const int xl = x > 0 ? x - 1 : (a.px ? W - 1 : 0);
    const int xr = x < W - 1 ? x + 1 : (a.px ? 0 : W - 1);
    const int yd = y > 0 ? y - 1 : (a.py ? H - 1 : 0);
    const int yu = y < H - 1 ? y + 1 : (a.py ? 0 : H - 1);
    T o = orig[(size_t)y * P + x];
    T upd = (o + ax * (at(xl, y) + at(xr, y)) + ay * (at(x, yd) + at(x, yu))) / den;
    out[(size_t)y * P + x] = upd;
    const unsigned long long k = mag_key((double)(upd - at(x, y)));
    m = k > m ? k : m;
  }
  if (a.direct) {
    // the sweep's max |upd - res| straight into its slot (a maximum does not depend on the order in
    // which the blocks arrive); for small grids, where this saves the fold's launch
    m = block_max_u64(m, lds);
    if (threadIdx.x == 0) key_max(&a.resid[c * a.stride + a.k], m);
  } else {
    put_max(a.part, c, m, lds);
  }
}

// Periodic duplicate endpoint copy after face diffusion (fb/core/grid.py:271-274).
template <class T>
__global__ void dup_copy_kernel(T* __restrict__ f, int rows, int cols, int pitch, int dup_x, int dup_y) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (dup_x && t < rows) f[(size_t)t * pitch + cols - 1] = f[(size_t)t * pitch];
  if (dup_y && t < cols) f[(size_t)(rows - 1) * pitch + t] = f[t];
}

// Wall velocity of a cell centre (fb/core/geometry.py:131-139), FP64.
__device__ __forceinline__ void wall_at(const MacGeom& g, int x, int y, double w, double pvx, double pvy, double& wu,
                                        double& wv) {
  double rx = cx_at(g, x) - pvx, ry = cy_at(g, y) - pvy;
  wu = (-w) * ry;
  wv = w * rx;
}

struct BcArgs {
  MacGeom g;
  double ux, uy;  // freestream (Python floats)
  int channel;
  double w, pvx, pvy;  // angular velocity (rad/s) and pivot
  const uint8_t* solid;
};

// apply_domain_boundaries then enforce_solid_faces on u faces, out of place
// (fb/core/grid.py:47-97).
template <class T>
__global__ void __launch_bounds__(256) bc_u_kernel(BcArgs b, const T* __restrict__ uin, T* __restrict__ uout) {
  const MacGeom& g = b.g;
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nx
  int j = blockIdx.y;
  if (i > g.nx) return;
  const size_t P = g.nx + 1;
  auto raw = [&](int jj, int ii) { return uin[(size_t)jj * P + ii]; };
  T val = raw(j, i);
  if (g.px) {
    if (i == 0 || i == g.nx) val = (T)0.5 * (raw(j, 0) + raw(j, g.nx));
  } else {
    if (i == 0) val = (T)b.ux;
    else if (i == g.nx) val = raw(j, g.nx - 1);
  }
  if (!g.py && (j == 0 || j == g.ny - 1)) val = b.channel ? (T)0 : (T)b.ux;
  // solid faces: a face touching a solid cell takes the face-averaged wall velocity
  bool left = i > 0 && b.solid[(size_t)j * g.nx + i - 1];
  bool right = i < g.nx && b.solid[(size_t)j * g.nx + i];
  if (left || right) {
    double wl, wr, tmp;
    if (i == 0) {
      wall_at(g, 0, j, b.w, b.pvx, b.pvy, wl, tmp);
      val = (T)wl;
    } else if (i == g.nx) {
      wall_at(g, g.nx - 1, j, b.w, b.pvx, b.pvy, wl, tmp);
      val = (T)wl;
    } else {
      wall_at(g, i - 1, j, b.w, b.pvx, b.pvy, wl, tmp);
      wall_at(g, i, j, b.w, b.pvx, b.pvy, wr, tmp);
      val = (T)(0.5 * (wl + wr));
    }
  }
  uout[(size_t)j * P + i] = val;
}

template <class T>
__global__ void __launch_bounds__(256) bc_v_kernel(BcArgs b, const T* __restrict__ vin, T* __restrict__ vout) {
  const MacGeom& g = b.g;
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nx-1
  int j = blockIdx.y;                             // 0..ny
  if (i >= g.nx) return;
  const size_t P = g.nx;
  auto raw = [&](int jj, int ii) { return vin[(size_t)jj * P + ii]; };
  // x-part (fb/core/grid.py:60-64)
  auto X = [&](int jj, int ii) -> T {
    if (!g.px) {
      if (ii == 0) return (T)b.uy;
      if (ii == g.nx - 1) return raw(jj, g.nx - 2);
    }
    return raw(jj, ii);
  };
  T val = X(j, i);
  if (j == 0 || j == g.ny) {
    if (g.py) val = (T)0.5 * (X(0, i) + X(g.ny, i));
    else if (b.channel) val = (T)0;
    else val = (T)b.uy;
  }
  bool below = j > 0 && b.solid[(size_t)(j - 1) * g.nx + i];
  bool above = j < g.ny && b.solid[(size_t)j * g.nx + i];
  if (below || above) {
    double tmp, wb, wa;
    if (j == 0) {
      wall_at(g, i, 0, b.w, b.pvx, b.pvy, tmp, wb);
      val = (T)wb;
    } else if (j == g.ny) {
      wall_at(g, i, g.ny - 1, b.w, b.pvx, b.pvy, tmp, wb);
      val = (T)wb;
    } else {
      wall_at(g, i, j - 1, b.w, b.pvx, b.pvy, tmp, wb);
      wall_at(g, i, j, b.w, b.pvx, b.pvy, tmp, wa);
      val = (T)(0.5 * (wb + wa));
    }
  }
  vout[(size_t)j * P + i] = val;
}

// rhs = fluid ? -div / max(dt, 1e-12) : 0 (fb/core/grid.py:148-150), dtype arithmetic, then FP64.
// ---------------------------------------------------------------- face-native transports

// sample_staggered_scalar (fb/core/interpolation.py:9-52) on a W x H field (row pitch W) whose
// point (0, 0) sits at cell-unit offset (ox, oy): FP64 weights, the result rounded to T. Periodic
// axes wrap with the cell count (a duplicated endpoint is never read).
template <class T>
__device__ __forceinline__ T sample_field(const MacGeom& g, const T* __restrict__ f, int W, int H, double ox,
                                          double oy, double px, double py) {
  double gx = (px - g.x0) / g.dx - ox;
  double gy = (py - g.y0) / g.dy - oy;
  gx = g.px ? pymod(gx, (double)g.nx) : fmin(fmax(gx, 0.0), (double)W - 1.0);
  gy = g.py ? pymod(gy, (double)g.ny) : fmin(fmax(gy, 0.0), (double)H - 1.0);
  const double fx = floor(gx), fy = floor(gy);
  int ix0 = (int)fx, iy0 = (int)fy;
  const double tx = gx - fx, ty = gy - fy;
  // np.mod can round a tiny negative argument up to the period itself: NumPy then reads index nx
  // (ny), which on a face field is the duplicated endpoint, and wraps the second index to 1; only
  // an index past the field's extent is clamped
  ix0 = min(max(ix0, 0), W - 1);
  iy0 = min(max(iy0, 0), H - 1);
  const int ix1 = g.px ? (ix0 + 1) % g.nx : min(ix0 + 1, W - 1);
  const int iy1 = g.py ? (iy0 + 1) % g.ny : min(iy0 + 1, H - 1);
  const double a = (double)f[(size_t)iy0 * W + ix0], b = (double)f[(size_t)iy0 * W + ix1];
  const double c = (double)f[(size_t)iy1 * W + ix0], d = (double)f[(size_t)iy1 * W + ix1];
  return (T)((1.0 - tx) * (1.0 - ty) * a + tx * (1.0 - ty) * b + (1.0 - tx) * ty * c + tx * ty * d);
}

// Face coordinates (fb/core/grid.py:306-320): u faces at (linspace(x0, x1, nx+1), cell-centre y),
// v faces at (cell-centre x, linspace(y0, y1, ny+1)).
__device__ __forceinline__ double ufx(const MacGeom& g, int i) { return linspace_at(i, g.nx + 1, g.x0, g.dx, g.x1); }
__device__ __forceinline__ double vfy(const MacGeom& g, int j) { return linspace_at(j, g.ny + 1, g.y0, g.dy, g.y1); }

// Face advection (fb/core/grid.py:323-376), both components in one launch over a
// (nx + 1) x (ny + 1) grid of threads. Without `orig`: each face takes fu/fv at its RK2 departure
// point through the cell velocity (cu, cv). With `orig` (the MacCormack correction, dt negated by
// the caller): the face's back-traced value b of fu/fv gives first + 0.5 (orig - b), clipped to
// orig's edge-replicated 3 x 3 bounds.
template <class T>
__global__ void __launch_bounds__(256) face_advect_kernel(MacGeom g, const T* __restrict__ cu, const T* __restrict__ cv,
                                                          const T* __restrict__ fu, const T* __restrict__ fv,
                                                          const T* __restrict__ ou, const T* __restrict__ ov,
                                                          double dt, T* __restrict__ out_u, T* __restrict__ out_v) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  auto bounded = [&](const T* orig, int W, int H, T c) {
    T lo = orig[(size_t)j * W + i], hi = lo;
    for (int oy = -1; oy <= 1; ++oy)
      for (int ox = -1; ox <= 1; ++ox) {
        const int jj = min(max(j + oy, 0), H - 1), ii = min(max(i + ox, 0), W - 1);
        const T q = orig[(size_t)jj * W + ii];
        lo = q < lo ? q : lo;
        hi = q > hi ? q : hi;
      }
    c = c > lo ? c : lo;
    return c < hi ? c : hi;
  };
  if (j < g.ny && i <= g.nx) {
    const int W = g.nx + 1;
    double qx, qy;
    backtrace(g, cu, cv, ufx(g, i), cy_at(g, j), dt, qx, qy);
    const T b = sample_field(g, fu, W, g.ny, 0.0, 0.5, qx, qy);
    const size_t k = (size_t)j * W + i;
    out_u[k] = ou ? bounded(ou, W, g.ny, fu[k] + (T)0.5 * (ou[k] - b)) : b;
  }
  if (j <= g.ny && i < g.nx) {
    const int W = g.nx;
    double qx, qy;
    backtrace(g, cu, cv, cx_at(g, i), vfy(g, j), dt, qx, qy);
    const T b = sample_field(g, fv, W, g.ny + 1, 0.5, 0.0, qx, qy);
    const size_t k = (size_t)j * W + i;
    out_v[k] = ov ? bounded(ov, W, g.ny + 1, fv[k] + (T)0.5 * (ov[k] - b)) : b;
  }
}

// centered_derivative (fb/core/grid.py:379-405) at point k of a line of n points spaced h, the
// values given by f(k): central differences with wrap (over n - 1 logical points when the last
// duplicates the first, whose derivative it takes), else numpy.gradient with second-order
// one-sided edges. Arithmetic in T with the spacing terms rounded to T, as NumPy does.
template <class T, class F>
__device__ __forceinline__ T cdiff(F&& f, int k, int n, double h, bool periodic, bool dup) {
  if (periodic) {
    const int L = dup ? n - 1 : n;
    if (k >= L) k = 0;
    const int kp = k + 1 == L ? 0 : k + 1, km = k == 0 ? L - 1 : k - 1;
    return (f(kp) - f(km)) / (T)(2.0 * h);
  }
  if (k == 0) return (T)(-1.5 / h) * f(0) + (T)(2.0 / h) * f(1) + (T)(-0.5 / h) * f(2);
  if (k == n - 1) return (T)(0.5 / h) * f(n - 3) + (T)(-2.0 / h) * f(n - 2) + (T)(1.5 / h) * f(n - 1);
  return (f(k + 1) - f(k - 1)) / (T)(2.0 * h);
}

// _cross_velocity_on_faces (fb/core/grid.py:408-431): v averaged onto u face (j, i) through the
// cells on either side, u onto v face (j, i).
template <class T>
__device__ __forceinline__ T v_on_u(const MacGeom& g, const T* __restrict__ v, int j, int i) {
  auto cell_v = [&](int ii) { return (T)0.5 * (v[(size_t)j * g.nx + ii] + v[(size_t)(j + 1) * g.nx + ii]); };
  if (i > 0 && i < g.nx) return (T)0.5 * (cell_v(i - 1) + cell_v(i));
  if (g.px) return (T)0.5 * (cell_v(g.nx - 1) + cell_v(0));
  return i == 0 ? cell_v(0) : cell_v(g.nx - 1);
}
template <class T>
__device__ __forceinline__ T u_on_v(const MacGeom& g, const T* __restrict__ u, int j, int i) {
  const size_t W = g.nx + 1;
  auto cell_u = [&](int jj) { return (T)0.5 * (u[(size_t)jj * W + i] + u[(size_t)jj * W + i + 1]); };
  if (j > 0 && j < g.ny) return (T)0.5 * (cell_u(j - 1) + cell_u(j));
  if (g.py) return (T)0.5 * (cell_u(g.ny - 1) + cell_u(0));
  return j == 0 ? cell_u(0) : cell_u(g.ny - 1);
}

// One skew-RK2 stage (fb/core/grid.py:446-503): out = base - c * S(su, sv), S the skew-symmetric
// convection (half the advective plus half the conservative form) of the faces (su, sv), c the
// stage's 0.5 dt or dt rounded to T. Boundaries are imposed by the caller.
template <class T>
__global__ void __launch_bounds__(256) skew_stage_kernel(MacGeom g, const T* __restrict__ su, const T* __restrict__ sv,
                                                         const T* __restrict__ bu, const T* __restrict__ bv, double c,
                                                         T* __restrict__ out_u, T* __restrict__ out_v) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  const bool px = g.px != 0, py = g.py != 0;
  const T ct = (T)c;
  if (j < g.ny && i <= g.nx) {
    const size_t W = g.nx + 1;
    auto U = [&](int jj, int ii) { return su[(size_t)jj * W + ii]; };
    const T uc = U(j, i), vu = v_on_u(g, sv, j, i);
    const T dudx = cdiff<T>([&](int k) { return U(j, k); }, i, g.nx + 1, g.dx, px, px);
    const T dudy = cdiff<T>([&](int k) { return U(k, i); }, j, g.ny, g.dy, py, false);
    const T adv = uc * dudx + vu * dudy;
    const T con = cdiff<T>([&](int k) { const T q = U(j, k); return q * q; }, i, g.nx + 1, g.dx, px, px) +
                  cdiff<T>([&](int k) { return v_on_u(g, sv, k, i) * U(k, i); }, j, g.ny, g.dy, py, false);
    const size_t k = (size_t)j * W + i;
    out_u[k] = bu[k] - ct * ((T)0.5 * (adv + con));
  }
  if (j <= g.ny && i < g.nx) {
    const size_t W = g.nx;
    auto V = [&](int jj, int ii) { return sv[(size_t)jj * W + ii]; };
    const T vc = V(j, i), uv = u_on_v(g, su, j, i);
    const T dvdx = cdiff<T>([&](int k) { return V(j, k); }, i, g.nx, g.dx, px, false);
    const T dvdy = cdiff<T>([&](int k) { return V(k, i); }, j, g.ny + 1, g.dy, py, py);
    const T adv = uv * dvdx + vc * dvdy;
    const T con = cdiff<T>([&](int k) { return u_on_v(g, su, j, k) * V(j, k); }, i, g.nx, g.dx, px, false) +
                  cdiff<T>([&](int k) { const T q = V(k, i); return q * q; }, j, g.ny + 1, g.dy, py, py);
    const size_t k = (size_t)j * W + i;
    out_v[k] = bv[k] - ct * ((T)0.5 * (adv + con));
  }
}

// skew_face_advection_rate (fb/core/grid.py:434-443): per-block maxima of the face rates, in T,
// into one order-preserving key.
template <class T>
__global__ void __launch_bounds__(256) skew_rate_kernel(MacGeom g, const T* __restrict__ u, const T* __restrict__ v,
                                                        unsigned long long* key) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  double r = 0.0;
  const T rdx = (T)g.dx, rdy = (T)g.dy;
  if (j < g.ny && i <= g.nx) {
    const T a = (T)fabs((double)u[(size_t)j * (g.nx + 1) + i]), b = (T)fabs((double)v_on_u(g, v, j, i));
    r = fmax(r, (double)(a / rdx + b / rdy));
  }
  if (j <= g.ny && i < g.nx) {
    const T a = (T)fabs((double)u_on_v(g, u, j, i)), b = (T)fabs((double)v[(size_t)j * g.nx + i]);
    r = fmax(r, (double)(a / rdx + b / rdy));
  }
  __shared__ unsigned long long lmax[4];
  const unsigned long long k = block_max_u64(dkey(r), lmax);
  if (threadIdx.x == 0) key_max(key, k);
}

template <class T>
__global__ void __launch_bounds__(256) rhs_kernel(MacGeom g, const T* __restrict__ u, const T* __restrict__ v,
                                                  const uint8_t* __restrict__ solid, double dt,
                                                  double* __restrict__ rhs) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= g.nx) return;
  size_t i = (size_t)y * g.nx + x;
  T du = u[(size_t)y * (g.nx + 1) + x + 1] - u[(size_t)y * (g.nx + 1) + x];
  T dv = v[(size_t)(y + 1) * g.nx + x] - v[(size_t)y * g.nx + x];
  T div = du / (T)g.dx + dv / (T)g.dy;
  T r = solid[i] ? (T)0 : (-div) / (T)fmax(dt, 1.0e-12);
  rhs[i] = (double)r;
}

// Pressure gradient (fb/core/grid.py:164-171), FP64 then rounded; out of place.
template <class T>
__global__ void __launch_bounds__(256) grad_u_kernel(MacGeom g, const T* __restrict__ uin, const double* __restrict__ p,
                                                     double dt, T* __restrict__ uout) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i > g.nx) return;
  const size_t P = g.nx + 1;
  const double* pr = p + (size_t)j * g.nx;
  double val = (double)uin[(size_t)j * P + i];
  if (i > 0 && i < g.nx) {
    val = val - dt * (pr[i] - pr[i - 1]) / g.dx;
  } else if (g.px) {
    // u[:, 0] -= dt (p[:, 0] - p[:, -1]) / dx; u[:, -1] = u[:, 0]
    val = (double)uin[(size_t)j * P] - dt * (pr[0] - pr[g.nx - 1]) / g.dx;
  }
  uout[(size_t)j * P + i] = (T)val;
}

template <class T>
__global__ void __launch_bounds__(256) grad_v_kernel(MacGeom g, const T* __restrict__ vin, const double* __restrict__ p,
                                                     double dt, T* __restrict__ vout) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= g.nx) return;
  const size_t P = g.nx;
  double val = (double)vin[(size_t)j * P + i];
  if (j > 0 && j < g.ny) {
    val = val - dt * (p[(size_t)j * g.nx + i] - p[(size_t)(j - 1) * g.nx + i]) / g.dy;
  } else if (g.py) {
    val = (double)vin[i] - dt * (p[i] - p[(size_t)(g.ny - 1) * g.nx + i]) / g.dy;
  }
  vout[(size_t)j * P + i] = (T)val;
}

// Reductions over the MAC state (keys: order-free max; counts):
// [0] max|u|,|v| faces  [1] max wall speed on solid cells  [2] max cell speed (dtype norm)
// [3] fluid divergence Linf  [4] solid-face leakage  [5] non-finite faces  [6] non-finite wall
struct StatArgs {
  MacGeom g;
  double w, pvx, pvy;
  const uint8_t* solid;
  unsigned long long* part;
};

// Partial slots: [0] max |face| over finite faces [1] max wall speed at solid cells [2] max
// cell-centred speed [3] max |div| over fluid cells [4] max solid leakage [5] non-finite faces
// [6] non-finite wall velocities.
template <class T>
__global__ void __launch_bounds__(256) mac_stats_kernel(StatArgs s, const T* __restrict__ u, const T* __restrict__ v) {
  __shared__ unsigned long long lds[4];
  const MacGeom& g = s.g;
  double face = 0.0;
  unsigned long long wsp = 0, csp = 0, dv = 0, leak = 0;
  unsigned long long nf = 0, nfw = 0;
  auto upd = [](unsigned long long& m, double val) {
    const unsigned long long k = mag_key(val);
    m = k > m ? k : m;
  };
  const size_t PU = g.nx + 1;
  const size_t total = (size_t)(g.nx + 1) * (g.ny + 1);
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
    const unsigned ui = (unsigned)idx, pu = (unsigned)(g.nx + 1);  // 32-bit index arithmetic
    const int y = (int)(ui / pu), x = (int)(ui - (unsigned)y * pu);
    if (y < g.ny) {
      T a = u[(size_t)y * PU + x];
      if (!isfinite((double)a)) nf++;
      else face = fmax(face, (double)fabs(a));
    }
    if (x < g.nx) {
      T b = v[(size_t)y * g.nx + x];
      if (!isfinite((double)b)) nf++;
      else face = fmax(face, (double)fabs(b));
    }
    if (x < g.nx && y < g.ny) {
      size_t i = (size_t)y * g.nx + x;
      bool sol = s.solid[i];
      double wu, wv;
      wall_at(g, x, y, s.w, s.pvx, s.pvy, wu, wv);
      if (!isfinite(wu) || !isfinite(wv)) nfw++;
      if (sol) upd(wsp, sqrt(wu * wu + wv * wv));
      T ul = u[(size_t)y * PU + x], ur = u[(size_t)y * PU + x + 1];
      T vb = v[i], va = v[(size_t)(y + 1) * g.nx + x];
      T cu = (T)0.5 * (ul + ur), cv = (T)0.5 * (vb + va);
      upd(csp, (double)sqrt(cu * cu + cv * cv));
      if (!sol) {
        T div = (ur - ul) / (T)g.dx + (va - vb) / (T)g.dy;
        upd(dv, (double)div);
      }
      // leakage on the right u face and the upper v face of this cell (interior interfaces)
      if (x + 1 < g.nx && (sol != (bool)s.solid[i + 1])) {
        double w1u, w1v, w2u, w2v;
        wall_at(g, x, y, s.w, s.pvx, s.pvy, w1u, w1v);
        wall_at(g, x + 1, y, s.w, s.pvx, s.pvy, w2u, w2v);
        upd(leak, (double)ur - 0.5 * (w1u + w2u));
      }
      if (y + 1 < g.ny && (sol != (bool)s.solid[i + g.nx])) {
        double w1u, w1v, w2u, w2v;
        wall_at(g, x, y, s.w, s.pvx, s.pvy, w1u, w1v);
        wall_at(g, x, y + 1, s.w, s.pvx, s.pvy, w2u, w2v);
        upd(leak, (double)va - 0.5 * (w1v + w2v));
      }
    }
  }
  put_max(s.part, 0, dkey(face), lds);
  put_max(s.part, 1, wsp, lds);
  put_max(s.part, 2, csp, lds);
  put_max(s.part, 3, dv, lds);
  put_max(s.part, 4, leak, lds);
  put_sum(s.part, 5, nf, lds);
  put_sum(s.part, 6, nfw, lds);
}

// Force on the foil per unit depth at unit density (build-defined; the reference computes none):
// on every interface between a fluid and a solid cell, the fluid cell's pressure pushes along the
// interface normal, and the tangential velocity of the fluid cell centre relative to the wall
// velocity of the solid cell, over half a cell, gives the viscous shear. FP64 sums per block, in a
// fixed order, into part[2 block], part[2 block + 1].
template <class T>
__global__ void __launch_bounds__(256) mac_force_kernel(MacGeom g, double nu, double w, double pvx, double pvy,
                                                        const uint8_t* __restrict__ solid,
                                                        const double* __restrict__ pres, const T* __restrict__ u,
                                                        const T* __restrict__ v, double* __restrict__ part) {
  __shared__ double lds[2][4];
  double fx = 0.0, fy = 0.0;
  const size_t PU = g.nx + 1;
  const size_t total = (size_t)g.nx * g.ny;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    const unsigned ui = (unsigned)i, unx = (unsigned)g.nx;  // 32-bit index arithmetic
    const int y = (int)(ui / unx), x = (int)(ui - (unsigned)y * unx);
    const bool s0 = solid[i] != 0;
    if (x + 1 < g.nx && s0 != (solid[i + 1] != 0)) {
      const size_t f = s0 ? i + 1 : i;
      const int xf = s0 ? x + 1 : x, xs = s0 ? x : x + 1;
      const double sign = s0 ? -1.0 : 1.0;  // normal from the fluid cell into the solid one
      fx += sign * pres[f] * g.dy;
      double wu, wv;
      wall_at(g, xs, y, w, pvx, pvy, wu, wv);
      const double vc = 0.5 * ((double)v[(size_t)y * g.nx + xf] + (double)v[(size_t)(y + 1) * g.nx + xf]);
      fy += nu * (vc - wv) / (0.5 * g.dx) * g.dy;
    }
    if (y + 1 < g.ny && s0 != (solid[i + g.nx] != 0)) {
      const size_t f = s0 ? i + g.nx : i;
      const int yf = s0 ? y + 1 : y, ys = s0 ? y : y + 1;
      const double sign = s0 ? -1.0 : 1.0;
      fy += sign * pres[f] * g.dx;
      double wu, wv;
      wall_at(g, x, ys, w, pvx, pvy, wu, wv);
      const double uc = 0.5 * ((double)u[(size_t)yf * PU + x] + (double)u[(size_t)yf * PU + x + 1]);
      fx += nu * (uc - wu) / (0.5 * g.dy) * g.dx;
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
__global__ void __launch_bounds__(256) sum_pairs_kernel(const double* __restrict__ part, int n,
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

// ---------------------------------------------------------------- host handle

int mac_geometry(MacGrid& M, const fb_pose& pose) {
  M.last_stats_valid = false;
  int rc = launch_cell_sdf(M.ctx->stream, make_cell_grid(M.dom), make_naca(M.foil, pose), M.sdf, M.solid);
  if (rc) return rc;
  M.angle = pose.angle_degrees;
  M.geom_valid = true;
  if ((rc = M.poisson.setup(M.ctx->stream, M.dom.nx, M.dom.ny, M.g.dx, M.g.dy, M.dom.periodic_x, M.dom.periodic_y,
                            M.solid)))
    return rc;
  M.poisson_ready = true;
  return FB_OK;
}

static int read_keys(MacGrid& M, int n) {
  FB_HIP(hipMemcpyAsync(M.host_keys, M.keys, n * sizeof(unsigned long long), hipMemcpyDeviceToHost, M.ctx->stream));
  FB_HIP(hipStreamSynchronize(M.ctx->stream));
  return FB_OK;
}

// Statistics of the current faces into keys[base .. base+6] (enqueued only).
template <class T>
static int launch_stats(MacGrid& M, double w, int base) {
  FB_HIP(hipMemsetAsync(M.keys + base, 0, 8 * sizeof(unsigned long long), M.ctx->stream));
  StatArgs s{M.g, w, M.foil.pivot_x, M.foil.pivot_y, M.solid, M.ctx->partials};
  const unsigned grid = reduction_blocks((size_t)(M.dom.nx + 1) * (M.dom.ny + 1));
  hipLaunchKernelGGL(mac_stats_kernel<T>, dim3(grid), dim3(256), 0, M.ctx->stream, s, (const T*)M.U[M.cur],
                     (const T*)M.V[M.cur]);
  FB_HIP(hipGetLastError());
  FinishSpec f;
  for (int q = 0; q < 7; ++q) f.add(base + q, q >= 5);
  return finish_partials(M.ctx, grid, f, M.keys);
}

// Decode keys[base ..] from the host copy.
static void stats_from_host(const MacGrid& M, int base, fb_mac_stats* st) {
  const unsigned long long* k = M.host_keys + base;
  auto val = [&](int j) { return k[j] ? dkey_inv(k[j]) : 0.0; };
  st->max_face_speed = val(0);
  st->max_wall_speed = val(1);
  st->max_cell_speed = val(2);
  st->divergence_linf = val(3);
  st->solid_leakage = val(4);
  st->nonfinite_faces = (int64_t)k[5];
  st->nonfinite_wall = (int64_t)k[6];
}

template <class T>
int mac_stats(MacGrid& M, double w, fb_mac_stats* st) {
  int rc = launch_stats<T>(M, w, 0);
  if (rc) return rc;
  if ((rc = read_keys(M, 7))) return rc;
  stats_from_host(M, 0, st);
  return FB_OK;
}

static double host_jacobi_scale(unsigned long long key) {
  const double m = key ? dkey_inv(key) : 0.0;
  return (m != m) ? m : std::max(m, 1.0);
}

// Jacobi implicit viscosity on two fields (cell planes, or u/v faces for PIC).
template <class T>
int diffuse(MacGrid& M, const T* in0, const T* in1, int W0, int H0, int P0, int W1, int H1, int P1,
                   double nu, double dt, double tol, int maxit, T** result, fb_sf_report* rep) {
  hipStream_t s = M.ctx->stream;
  const int px = M.dom.periodic_x, py = M.dom.periodic_y;
  int rc;
  if (nu <= 0.0) {
    result[0] = (T*)in0;
    result[1] = (T*)in1;
    rep->viscosity_iterations = 0;
    rep->viscosity_residual = 0.0;
    rep->viscosity_converged = 1;
    return FB_OK;
  }
  if (maxit + 2 > M.resid_cap) {
    if (M.resid) (void)dev_free(M.resid);
    if (M.host_resid) (void)host_free(M.host_resid);
    M.resid_cap = maxit + 2;
    FB_HIP(dev_alloc((void**)&M.resid, 2 * M.resid_cap * sizeof(unsigned long long)));
    FB_HIP(host_alloc((void**)&M.host_resid, 2 * M.resid_cap * sizeof(unsigned long long)));
  }
  FB_HIP(hipMemsetAsync(M.resid, 0, 2 * M.resid_cap * sizeof(unsigned long long), s));
  // scale = max(max |original|, 1), kept on the device (keys 0 and 1) and read with the residuals
  FB_HIP(hipMemsetAsync(M.keys, 0, 2 * sizeof(unsigned long long), s));
  {
    const unsigned nb = reduction_blocks(std::max((size_t)H0 * P0, (size_t)H1 * P1));
    hipLaunchKernelGGL(absmax_kernel<T>, dim3(nb), dim3(256), 0, s, (size_t)H0 * P0, in0, M.ctx->partials, 0);
    hipLaunchKernelGGL(absmax_kernel<T>, dim3(nb), dim3(256), 0, s, (size_t)H1 * P1, in1, M.ctx->partials, 1);
    FB_HIP(hipGetLastError());
    FinishSpec f;
    f.add(0, false);
    f.add(1, false);
    if ((rc = finish_partials(M.ctx, nb, f, M.keys))) return rc;
  }
  JacobiArgs<T> a{};
  a.orig[0] = in0;
  a.orig[1] = in1;
  a.W[0] = W0;
  a.H[0] = H0;
  a.P[0] = P0;
  a.W[1] = W1;
  a.H[1] = H1;
  a.P[1] = P1;
  a.px = px;
  a.py = py;
  a.ax = nu * dt / (M.g.dx * M.g.dx);
  a.ay = nu * dt / (M.g.dy * M.g.dy);
  a.den = 1.0 + 2.0 * a.ax + 2.0 * a.ay;
  a.scale_keys = M.keys;
  a.tol = tol;
  a.resid = M.resid;
  a.stride = M.resid_cap;
  a.part = M.ctx->partials;
  T* buf[2][2] = {{(T*)M.DU[1], (T*)M.DU[2]}, {(T*)M.DV[1], (T*)M.DV[2]}};
  dim3 grid(reduction_blocks(std::max((size_t)W0 * H0, (size_t)W1 * H1)), 1, 2);
  // same-address maxima serialise: only small grids fold inside the sweep
  a.direct = grid.x <= 256 ? 1 : 0;
  int done[2] = {0, 0};
  double fres[2] = {INFINITY, INFINITY};
  int k = 1;
  // the first read comes after about as many sweeps as the previous solve needed
  int chunk = std::min(std::max(M.last_sweeps + 1, 4), 64);
  while (k <= maxit && !(done[0] && done[1])) {
    int kend = std::min(maxit, k + chunk - 1);
    for (int kk = k; kk <= kend; ++kk) {
      for (int c = 0; c < 2; ++c) {
        a.in[c] = (kk == 1) ? a.orig[c] : buf[c][(kk - 1) % 2];
        a.out[c] = buf[c][kk % 2];
      }
      a.k = kk;
      hipLaunchKernelGGL(jacobi_sweep_kernel<T>, grid, dim3(256), 0, s, a);
      FB_HIP(hipGetLastError());
      if (!a.direct) {
        FinishSpec f;
        f.add(kk, false);
        f.add(M.resid_cap + kk, false);
        if ((rc = finish_partials(M.ctx, grid.x, f, M.resid))) return rc;
      }
    }
    FB_HIP(hipMemcpyAsync(M.host_resid, M.resid, 2 * M.resid_cap * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
    FB_HIP(hipMemcpyAsync(M.host_keys, M.keys, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    FB_HIP(hipStreamSynchronize(s));
    const double scale[2] = {host_jacobi_scale(M.host_keys[0]), host_jacobi_scale(M.host_keys[1])};
    for (int c = 0; c < 2; ++c) {
      if (done[c]) continue;
      for (int kk = k; kk <= kend; ++kk) {
        double r = dkey_inv(M.host_resid[c * M.resid_cap + kk]) / scale[c];
        fres[c] = r;
        if (r <= tol) {
          done[c] = kk;
          break;
        }
      }
    }
    k = kend + 1;
    chunk = std::min(chunk * 2, 64);
  }
  int its[2];
  for (int c = 0; c < 2; ++c) {
    its[c] = done[c] ? done[c] : maxit;
    result[c] = buf[c][its[c] % 2];
  }
  M.last_sweeps = std::max(its[0], its[1]);
  rep->viscosity_iterations = std::max(its[0], its[1]);
  rep->viscosity_residual = std::max(fres[0], fres[1]);
  rep->viscosity_converged = (fres[0] <= tol && fres[1] <= tol) ? 1 : 0;
  return FB_OK;
}

// Extrapolated warm start: p <- 2 p - prev (when `extrapolate`), prev <- the old p.
__global__ void __launch_bounds__(256) extrapolate_kernel(size_t N, double* __restrict__ p, double* __restrict__ prev,
                                                          int extrapolate) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const double t = p[i];
  if (extrapolate) p[i] = 2.0 * t - prev[i];
  prev[i] = t;
}

// A warm-started projection that ends with any non-zero status leaves no usable initial guess:
// the pressure is zeroed and the extrapolation history dropped.
static int warm_reset_on_failure(MacGrid& M, const fb_sf_params& p, const fb_sf_report* rep) {
  if (p.warm_pressure && rep->status != 0) {
    FB_HIP(hipMemsetAsync(M.pres, 0, M.N * sizeof(double), M.ctx->stream));
    M.pres_solves = 0;
  }
  return FB_OK;
}

template <class T>
int project(MacGrid& M, const fb_sf_params& p, double dt, fb_sf_report* rep) {
  // The reference checks the CFL guard and finiteness before solving (fb/solvers/stable_fluids.py:
  // 256-330). Here the statistics are taken before the solve and read with everything else at the
  // single synchronisation at the end; the status follows the reference's order of checks. The
  // faces are modified either way, and the callers roll back on any non-zero status.
  hipStream_t s = M.ctx->stream;
  int rc;
  if ((rc = launch_stats<T>(M, p.wall_omega_rad, 0))) return rc;
  BcArgs b{M.g, p.freestream_x, p.freestream_y, p.channel_walls, p.wall_omega_rad, M.foil.pivot_x, M.foil.pivot_y,
           M.solid};
  const int a = M.cur, o = 1 - M.cur;
  dim3 gu((M.dom.nx + 1 + 255) / 256, M.dom.ny), gv((M.dom.nx + 255) / 256, M.dom.ny + 1),
      gc((M.dom.nx + 255) / 256, M.dom.ny);
  hipLaunchKernelGGL(bc_u_kernel<T>, gu, dim3(256), 0, s, b, (const T*)M.U[a], (T*)M.U[o]);
  hipLaunchKernelGGL(bc_v_kernel<T>, gv, dim3(256), 0, s, b, (const T*)M.V[a], (T*)M.V[o]);
  hipLaunchKernelGGL(rhs_kernel<T>, gc, dim3(256), 0, s, M.g, (const T*)M.U[o], (const T*)M.V[o], M.solid, dt,
                     M.rhs);
  FB_HIP(hipGetLastError());
  PoissonReport pr;
  const bool hooked = M.hook.fn != nullptr;
  const int mode = p.mixed_poisson | (p.warm_pressure ? kWarmStart : 0);
  if (p.warm_pressure == 2) {
    // initial guess 2 p_n - p_(n-1) once two solves are on record, else p_n
    if (!M.pres_prev) {
      FB_HIP(dev_alloc((void**)&M.pres_prev, M.N * sizeof(double)));
      M.pres_solves = 0;
    }
    hipLaunchKernelGGL(extrapolate_kernel, dim3((unsigned)((M.N + 255) / 256)), dim3(256), 0, s, M.N, M.pres,
                       M.pres_prev, M.pres_solves >= 2 ? 1 : 0);
    FB_HIP(hipGetLastError());
  }
  if (hooked) rc = M.hook.fn(M.hook.user, M, p.tolerance, p.max_iterations, mode, &pr);
  else rc = M.poisson.solve(M.rhs, M.pres, p.tolerance, p.max_iterations, mode, &pr);
  if (rc) return rc;
  M.pres_solves++;
  hipLaunchKernelGGL(grad_u_kernel<T>, gu, dim3(256), 0, s, M.g, (const T*)M.U[o], M.pres, dt, (T*)M.U[a]);
  hipLaunchKernelGGL(grad_v_kernel<T>, gv, dim3(256), 0, s, M.g, (const T*)M.V[o], M.pres, dt, (T*)M.V[a]);
  hipLaunchKernelGGL(bc_u_kernel<T>, gu, dim3(256), 0, s, b, (const T*)M.U[a], (T*)M.U[o]);
  hipLaunchKernelGGL(bc_v_kernel<T>, gv, dim3(256), 0, s, b, (const T*)M.V[a], (T*)M.V[o]);
  FB_HIP(hipGetLastError());
  M.cur = o;
  if ((rc = launch_stats<T>(M, p.wall_omega_rad, 8))) return rc;
  FB_HIP(hipMemcpyAsync(M.host_keys, M.keys, 16 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  FB_HIP(hipStreamSynchronize(s));
  if (!hooked) M.poisson.finish(&pr);
  fb_mac_stats pre{}, post{};
  stats_from_host(M, 0, &pre);
  stats_from_host(M, 8, &post);
  if (pre.nonfinite_faces || pre.nonfinite_wall) {
    rep->status = FB_SF_NONFINITE_PROJECTION_INPUT;
    return warm_reset_on_failure(M, p, rep);
  }
  const double pc = std::max(pre.max_face_speed, pre.max_wall_speed) * dt / std::min(M.g.dx, M.g.dy);
  rep->projection_cfl = pc;
  if (pc > p.cfl_limit) {
    rep->status = FB_SF_PROJECTION_CFL;
    return warm_reset_on_failure(M, p, rep);
  }
  rep->pressure_iterations = pr.iterations;
  rep->pressure_residual = pr.relative_residual;
  rep->pressure_info = pr.info;
  rep->pressure_refinements = pr.refinements;
  if (pr.failure) {
    rep->status = pr.failure == 1 ? FB_SF_NONFINITE_RHS
                  : pr.failure == 2 ? FB_SF_RHS_RANGE
                  : pr.failure == 3 ? FB_SF_NONFINITE_ITERATE
                                    : FB_SF_CG_BREAKDOWN;
    return warm_reset_on_failure(M, p, rep);
  }
  rep->pressure_converged =
      (pr.info == 0 && pr.relative_residual <= p.tolerance * (1.0 + 1.0e-6)) ? 1 : 0;
  if (!rep->pressure_converged) {
    rep->status = FB_SF_PROJECTION_NOT_CONVERGED;
    return warm_reset_on_failure(M, p, rep);
  }
  if (post.nonfinite_faces) rep->status = FB_SF_NONFINITE_PROJECTION_OUTPUT;
  // the statistics of the projected faces, for the caller's postconditions
  M.last_stats = post;
  M.last_stats_omega = p.wall_omega_rad;
  M.last_stats_valid = true;
  return warm_reset_on_failure(M, p, rep);
}

// Face-native transports with face diffusion (fb/solvers/stable_fluids.py:431-460): skew-RK2
// (face_transport 2) or semi-Lagrangian / MacCormack face advection (1), then implicit
// viscosity on each face component; the faces end in U[cur], V[cur].
template <class T>
static int face_transport_step(MacGrid& M, const fb_sf_params& p, fb_sf_report* rep) {
  hipStream_t s = M.ctx->stream;
  int rc;
  const int nx = M.dom.nx, ny = M.dom.ny, o = 1 - M.cur;
  T* u = (T*)M.U[M.cur];
  T* v = (T*)M.V[M.cur];
  T* ou = (T*)M.U[o];
  T* ov = (T*)M.V[o];
  T* tu = (T*)M.DU[0];
  T* tv = (T*)M.DV[0];
  const dim3 gf((nx + 1 + 255) / 256, ny + 1);
  if (p.face_transport == 2) {
    // boundaries without channel walls (grid.py:495, 501) and the mask of the step's start
    BcArgs b{M.g, p.freestream_x, p.freestream_y, 0, p.wall_omega_rad, M.foil.pivot_x, M.foil.pivot_y, M.solid};
    const dim3 gu((nx + 1 + 255) / 256, ny), gv((nx + 255) / 256, ny + 1);
    hipLaunchKernelGGL(skew_stage_kernel<T>, gf, dim3(256), 0, s, M.g, (const T*)u, (const T*)v, (const T*)u,
                       (const T*)v, 0.5 * p.dt, tu, tv);
    hipLaunchKernelGGL(bc_u_kernel<T>, gu, dim3(256), 0, s, b, (const T*)tu, ou);
    hipLaunchKernelGGL(bc_v_kernel<T>, gv, dim3(256), 0, s, b, (const T*)tv, ov);
    hipLaunchKernelGGL(skew_stage_kernel<T>, gf, dim3(256), 0, s, M.g, (const T*)ou, (const T*)ov, (const T*)u,
                       (const T*)v, p.dt, tu, tv);
    hipLaunchKernelGGL(bc_u_kernel<T>, gu, dim3(256), 0, s, b, (const T*)tu, ou);
    hipLaunchKernelGGL(bc_v_kernel<T>, gv, dim3(256), 0, s, b, (const T*)tv, ov);
  } else {
    T* cu = (T*)M.CU[0];
    T* cv = (T*)M.CV[0];
    if ((rc = launch_faces_to_cell<T>(M, u, v, cu, cv))) return rc;
    hipLaunchKernelGGL(face_advect_kernel<T>, gf, dim3(256), 0, s, M.g, (const T*)cu, (const T*)cv, (const T*)u,
                       (const T*)v, (const T*)nullptr, (const T*)nullptr, p.dt, p.maccormack ? tu : ou,
                       p.maccormack ? tv : ov);
    if (p.maccormack) {
      T* fcu = (T*)M.CU[1];
      T* fcv = (T*)M.CV[1];
      if ((rc = launch_faces_to_cell<T>(M, tu, tv, fcu, fcv))) return rc;
      hipLaunchKernelGGL(face_advect_kernel<T>, gf, dim3(256), 0, s, M.g, (const T*)fcu, (const T*)fcv,
                         (const T*)tu, (const T*)tv, (const T*)u, (const T*)v, -p.dt, ou, ov);
    }
  }
  FB_HIP(hipGetLastError());
  // implicit viscosity on the faces (grid.py:278-303): Jacobi over the logical faces, the
  // duplicated periodic endpoint copied from the first
  T* res[2];
  if ((rc = diffuse<T>(M, ou, ov, nx + 1 - M.g.px, ny, nx + 1, nx, ny + 1 - M.g.py, nx, p.viscosity, p.dt,
                       p.tolerance, p.max_iterations, res, rep)))
    return rc;
  if (p.viscosity > 0.0) {
    if ((rc = launch_dup_copy<T>(M, res[0], ny, nx + 1, nx + 1, M.g.px, 0))) return rc;
    if ((rc = launch_dup_copy<T>(M, res[1], ny + 1, nx, nx, 0, M.g.py))) return rc;
  }
  FB_HIP(hipMemcpyAsync(u, res[0], M.NU * M.sz, hipMemcpyDeviceToDevice, s));
  FB_HIP(hipMemcpyAsync(v, res[1], M.NV * M.sz, hipMemcpyDeviceToDevice, s));
  return FB_OK;
}

template <class T>
static int substep(MacGrid& M, const fb_sf_params& p, const fb_pose& pose, fb_sf_report* rep) {
  hipStream_t s = M.ctx->stream;
  M.last_stats_valid = false;
  int rc;
  const int nx = M.dom.nx, ny = M.dom.ny;
  if (p.face_transport) {
    if ((rc = face_transport_step<T>(M, p, rep))) return rc;
    if (!rep->viscosity_converged) {
      rep->status = FB_SF_VISCOSITY_NOT_CONVERGED;
      return FB_OK;
    }
    if (!M.geom_valid || pose.angle_degrees != M.angle) {
      if ((rc = mac_geometry(M, pose))) return rc;
    }
    return project<T>(M, p, p.dt, rep);
  }
  dim3 gc((nx + 255) / 256, ny);
  T* cu = (T*)M.CU[0];
  T* cv = (T*)M.CV[0];
  hipLaunchKernelGGL(faces_to_cell_kernel<T>, gc, dim3(256), 0, s, nx, ny, (const T*)M.U[M.cur],
                     (const T*)M.V[M.cur], cu, cv);
  hipLaunchKernelGGL(advect_first_kernel<T>, gc, dim3(256), 0, s, M.g, (const T*)cu, (const T*)cv, p.dt,
                     (T*)M.CU[1], (T*)M.CV[1]);
  FB_HIP(hipGetLastError());
  const T* au = (const T*)M.CU[1];
  const T* av = (const T*)M.CV[1];
  if (p.maccormack) {
    hipLaunchKernelGGL(advect_correct_kernel<T>, gc, dim3(256), 0, s, M.g, (const T*)cu, (const T*)cv,
                       (const T*)M.CU[1], (const T*)M.CV[1], p.dt, (T*)M.CU[2], (T*)M.CV[2]);
    FB_HIP(hipGetLastError());
    au = (const T*)M.CU[2];
    av = (const T*)M.CV[2];
  }
  T* res[2];
  if ((rc = diffuse<T>(M, au, av, nx, ny, nx, nx, ny, nx, p.viscosity, p.dt, p.tolerance, p.max_iterations, res, rep)))
    return rc;
  dim3 gf((nx + 1 + 255) / 256, ny + 1);
  hipLaunchKernelGGL(cell_to_faces_kernel<T>, gf, dim3(256), 0, s, nx, ny, (const T*)res[0], (const T*)res[1],
                     (T*)M.U[M.cur], (T*)M.V[M.cur]);
  FB_HIP(hipGetLastError());
  if (!rep->viscosity_converged) {
    rep->status = FB_SF_VISCOSITY_NOT_CONVERGED;
    return FB_OK;
  }
  if (!M.geom_valid || pose.angle_degrees != M.angle) {
    if ((rc = mac_geometry(M, pose))) return rc;
  }
  return project<T>(M, p, p.dt, rep);
}

// Launch helpers of the grid kernels shared with the PIC/FLIP path.
template <class T>
int launch_faces_to_cell(MacGrid& M, const T* u, const T* v, T* cu, T* cv) {
  dim3 gc((M.dom.nx + 255) / 256, M.dom.ny);
  hipLaunchKernelGGL(faces_to_cell_kernel<T>, gc, dim3(256), 0, M.ctx->stream, M.dom.nx, M.dom.ny, u, v, cu, cv);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

template <class T>
int launch_cell_to_faces(MacGrid& M, const T* cu, const T* cv, T* u, T* v) {
  dim3 gf((M.dom.nx + 1 + 255) / 256, M.dom.ny + 1);
  hipLaunchKernelGGL(cell_to_faces_kernel<T>, gf, dim3(256), 0, M.ctx->stream, M.dom.nx, M.dom.ny, cu, cv, u, v);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

template <class T>
int launch_dup_copy(MacGrid& M, T* f, int rows, int cols, int pitch, int dup_x, int dup_y) {
  if (!dup_x && !dup_y) return FB_OK;
  int n = std::max(rows, cols);
  hipLaunchKernelGGL(dup_copy_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, M.ctx->stream, f, rows, cols, pitch,
                     dup_x, dup_y);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

template int launch_faces_to_cell<double>(MacGrid&, const double*, const double*, double*, double*);
template int launch_faces_to_cell<float>(MacGrid&, const float*, const float*, float*, float*);
template int launch_cell_to_faces<double>(MacGrid&, const double*, const double*, double*, double*);
template int launch_cell_to_faces<float>(MacGrid&, const float*, const float*, float*, float*);
template int launch_dup_copy<double>(MacGrid&, double*, int, int, int, int, int);
template int launch_dup_copy<float>(MacGrid&, float*, int, int, int, int, int);
template int diffuse<double>(MacGrid&, const double*, const double*, int, int, int, int, int, int, double, double,
                             double, int, double**, fb_sf_report*);
template int diffuse<float>(MacGrid&, const float*, const float*, int, int, int, int, int, int, double, double, double,
                            int, float**, fb_sf_report*);
template int project<double>(MacGrid&, const fb_sf_params&, double, fb_sf_report*);
template int project<float>(MacGrid&, const fb_sf_params&, double, fb_sf_report*);
template int mac_stats<double>(MacGrid&, double, fb_mac_stats*);
template int mac_stats<float>(MacGrid&, double, fb_mac_stats*);

void free_mac(MacGrid* M) {
  for (void* ptr : {M->U[0], M->U[1], M->V[0], M->V[1], M->CU[0], M->CU[1], M->CU[2], M->CU[3], M->CV[0], M->CV[1],
                    M->CV[2], M->CV[3], M->DU[0], M->DU[1], M->DU[2], M->DV[0], M->DV[1], M->DV[2], M->CKU, M->CKV,
                    (void*)M->sdf, (void*)M->solid, (void*)M->ck_solid, (void*)M->rhs, (void*)M->pres,
                    (void*)M->pres_prev, (void*)M->keys, (void*)M->skew_key, (void*)M->resid})
    if (ptr) (void)dev_free(ptr);
  if (M->host_keys) (void)host_free(M->host_keys);
  if (M->host_resid) (void)host_free(M->host_resid);
  delete M;
}

}  // namespace fb

using fb::MacGrid;

#define MAC_DISPATCH(M, fn, ...) \
  ((M)->prec == FB_F64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

namespace fb {

int mac_substep(MacGrid& M, const fb_sf_params& p, const fb_pose& pose, fb_sf_report* rep) {
  *rep = fb_sf_report{};
  return MAC_DISPATCH(&M, substep, M, p, pose, rep);
}

void mac_bind(MacGrid& M, Ctx* c) {
  M.ctx = c;
  M.poisson.stream = c->stream;
}

void mac_use_graphs(MacGrid& M, bool on) { M.poisson.set_use_graphs(on); }

}  // namespace fb

extern "C" {

int fb_mac_create(fb_ctx* ctx, const fb_domain* dom, const fb_foil* foil, int precision, fb_mac** out) {
  FB_REQUIRE(ctx && dom && foil && out, FB_ERR_ARG, "NULL argument");
  MacGrid* M = nullptr;
  int rc = fb::create_mac(reinterpret_cast<fb::Ctx*>(ctx), dom, foil, precision, &M);
  if (rc) return rc;
  *out = reinterpret_cast<fb_mac*>(M);
  return FB_OK;
}

}  // extern "C"

namespace fb {

int create_mac(Ctx* c, const fb_domain* dom, const fb_foil* foil, int precision, MacGrid** out) {
  FB_REQUIRE(dom->nx >= 4 && dom->ny >= 4, FB_ERR_ARG, "each resolution axis must contain at least four cells");
  FB_REQUIRE(precision == FB_F64 || precision == FB_F32, FB_ERR_ARG, "MAC storage must be FB_F64 or FB_F32");
  FB_HIP(hipSetDevice(c->device));
  auto* M = new MacGrid();
  M->ctx = c;
  M->dom = *dom;
  M->foil = *foil;
  M->g = make_geom(*dom);
  M->prec = precision;
  M->sz = precision == FB_F64 ? 8 : 4;
  M->N = (size_t)dom->nx * dom->ny;
  M->NU = (size_t)dom->ny * (dom->nx + 1);
  M->NV = (size_t)(dom->ny + 1) * dom->nx;
  const size_t NF = std::max(M->NU, M->NV);
  hipError_t e = hipSuccess;
  auto A = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = dev_alloc(p, bytes);
  };
  for (int k = 0; k < 2; ++k) {
    A(&M->U[k], M->NU * M->sz);
    A(&M->V[k], M->NV * M->sz);
  }
  for (int k = 0; k < 4; ++k) {
    A(&M->CU[k], M->N * M->sz);
    A(&M->CV[k], M->N * M->sz);
  }
  for (int k = 0; k < 3; ++k) {
    A(&M->DU[k], NF * M->sz);
    A(&M->DV[k], NF * M->sz);
  }
  A(&M->CKU, M->NU * M->sz);
  A(&M->CKV, M->NV * M->sz);
  A((void**)&M->sdf, M->N * sizeof(double));
  A((void**)&M->solid, M->N);
  A((void**)&M->ck_solid, M->N);
  A((void**)&M->rhs, M->N * sizeof(double));
  A((void**)&M->pres, M->N * sizeof(double));
  // the pressure is the initial guess of a warm-started solve: zero before the first one
  if (e == hipSuccess) e = hipMemsetAsync(M->pres, 0, M->N * sizeof(double), c->stream);
  A((void**)&M->keys, 16 * sizeof(unsigned long long));
  A((void**)&M->skew_key, sizeof(unsigned long long));
  if (e == hipSuccess) e = host_alloc((void**)&M->host_keys, 16 * sizeof(unsigned long long));
  if (e != hipSuccess) {
    set_error("device allocation failed: %s", hipGetErrorString(e));
    free_mac(M);
    return FB_ERR_NOMEM;
  }
  *out = M;
  return FB_OK;
}

}  // namespace fb

extern "C" {

int fb_mac_destroy(fb_mac* mac) {
  if (!mac) return FB_OK;
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_HIP(hipSetDevice(M->ctx->device));
  FB_HIP(hipStreamSynchronize(M->ctx->stream));
  fb::free_mac(M);
  return FB_OK;
}

int fb_mac_set_faces(fb_mac* mac, const void* u_host, const void* v_host) {
  FB_REQUIRE(mac && u_host && v_host, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  M->last_stats_valid = false;
  FB_HIP(hipSetDevice(M->ctx->device));
  FB_HIP(hipMemcpyAsync(M->U[M->cur], u_host, M->NU * M->sz, hipMemcpyHostToDevice, M->ctx->stream));
  FB_HIP(hipMemcpyAsync(M->V[M->cur], v_host, M->NV * M->sz, hipMemcpyHostToDevice, M->ctx->stream));
  FB_HIP(hipStreamSynchronize(M->ctx->stream));
  return FB_OK;
}

int fb_mac_get_faces(fb_mac* mac, void* u_host, void* v_host) {
  FB_REQUIRE(mac, FB_ERR_ARG, "NULL handle");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_HIP(hipSetDevice(M->ctx->device));
  if (u_host) FB_HIP(hipMemcpyAsync(u_host, M->U[M->cur], M->NU * M->sz, hipMemcpyDeviceToHost, M->ctx->stream));
  if (v_host) FB_HIP(hipMemcpyAsync(v_host, M->V[M->cur], M->NV * M->sz, hipMemcpyDeviceToHost, M->ctx->stream));
  FB_HIP(hipStreamSynchronize(M->ctx->stream));
  return FB_OK;
}

int fb_mac_set_pose(fb_mac* mac, const fb_pose* pose) {
  FB_REQUIRE(mac && pose, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_HIP(hipSetDevice(M->ctx->device));
  return fb::mac_geometry(*M, *pose);
}

int fb_mac_get_geometry(fb_mac* mac, double* sdf_host, uint8_t* solid_host) {
  FB_REQUIRE(mac, FB_ERR_ARG, "NULL handle");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_REQUIRE(M->geom_valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(M->ctx->device));
  if (sdf_host) FB_HIP(hipMemcpyAsync(sdf_host, M->sdf, M->N * sizeof(double), hipMemcpyDeviceToHost, M->ctx->stream));
  if (solid_host) FB_HIP(hipMemcpyAsync(solid_host, M->solid, M->N, hipMemcpyDeviceToHost, M->ctx->stream));
  FB_HIP(hipStreamSynchronize(M->ctx->stream));
  return FB_OK;
}

int fb_mac_get_pressure(fb_mac* mac, double* p_host) {
  FB_REQUIRE(mac && p_host, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_HIP(hipSetDevice(M->ctx->device));
  FB_HIP(hipMemcpyAsync(p_host, M->pres, M->N * sizeof(double), hipMemcpyDeviceToHost, M->ctx->stream));
  FB_HIP(hipStreamSynchronize(M->ctx->stream));
  return FB_OK;
}

int fb_mac_set_pressure(fb_mac* mac, const double* p_host) {
  FB_REQUIRE(mac && p_host, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_HIP(hipSetDevice(M->ctx->device));
  FB_HIP(hipMemcpyAsync(M->pres, p_host, M->N * sizeof(double), hipMemcpyHostToDevice, M->ctx->stream));
  FB_HIP(hipStreamSynchronize(M->ctx->stream));
  M->pres_solves = 1;  // an imported pressure starts a new extrapolation history
  return FB_OK;
}

int fb_mac_skew_rate(fb_mac* mac, double* rate) {
  FB_REQUIRE(mac && rate, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_HIP(hipSetDevice(M->ctx->device));
  hipStream_t s = M->ctx->stream;
  const dim3 gf((M->dom.nx + 1 + 255) / 256, M->dom.ny + 1);
  FB_HIP(hipMemsetAsync(M->skew_key, 0, sizeof(unsigned long long), s));
  if (M->prec == FB_F64)
    hipLaunchKernelGGL(fb::skew_rate_kernel<double>, gf, dim3(256), 0, s, M->g, (const double*)M->U[M->cur],
                       (const double*)M->V[M->cur], M->skew_key);
  else
    hipLaunchKernelGGL(fb::skew_rate_kernel<float>, gf, dim3(256), 0, s, M->g, (const float*)M->U[M->cur],
                       (const float*)M->V[M->cur], M->skew_key);
  FB_HIP(hipGetLastError());
  unsigned long long k = 0;
  FB_HIP(hipMemcpyAsync(&k, M->skew_key, sizeof(k), hipMemcpyDeviceToHost, s));
  FB_HIP(hipStreamSynchronize(s));
  *rate = k ? fb::dkey_inv(k) : 0.0;
  return FB_OK;
}

int fb_mac_forces(fb_mac* mac, double viscosity, double wall_omega_rad, double* force) {
  FB_REQUIRE(mac && force, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_REQUIRE(M->geom_valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(M->ctx->device));
  hipStream_t s = M->ctx->stream;
  const unsigned nb = fb::reduction_blocks(M->N);
  double* part = reinterpret_cast<double*>(M->ctx->partials);  // 2 doubles per block fit the slots
  double* out = part + 2 * (size_t)fb::kPartialBlocks;
  if (M->prec == FB_F64)
    hipLaunchKernelGGL(fb::mac_force_kernel<double>, dim3(nb), dim3(256), 0, s, M->g, viscosity, wall_omega_rad,
                       M->foil.pivot_x, M->foil.pivot_y, M->solid, M->pres, (const double*)M->U[M->cur],
                       (const double*)M->V[M->cur], part);
  else
    hipLaunchKernelGGL(fb::mac_force_kernel<float>, dim3(nb), dim3(256), 0, s, M->g, viscosity, wall_omega_rad,
                       M->foil.pivot_x, M->foil.pivot_y, M->solid, M->pres, (const float*)M->U[M->cur],
                       (const float*)M->V[M->cur], part);
  hipLaunchKernelGGL(fb::sum_pairs_kernel, dim3(1), dim3(256), 0, s, part, (int)nb, out);
  FB_HIP(hipGetLastError());
  FB_HIP(hipMemcpyAsync(M->host_keys, out, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  FB_HIP(hipStreamSynchronize(s));
  memcpy(force, M->host_keys, 2 * sizeof(double));
  return FB_OK;
}

int fb_mac_checkpoint(fb_mac* mac) {
  FB_REQUIRE(mac, FB_ERR_ARG, "NULL handle");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_HIP(hipSetDevice(M->ctx->device));
  hipStream_t s = M->ctx->stream;
  FB_HIP(hipMemcpyAsync(M->CKU, M->U[M->cur], M->NU * M->sz, hipMemcpyDeviceToDevice, s));
  FB_HIP(hipMemcpyAsync(M->CKV, M->V[M->cur], M->NV * M->sz, hipMemcpyDeviceToDevice, s));
  if (M->geom_valid) FB_HIP(hipMemcpyAsync(M->ck_solid, M->solid, M->N, hipMemcpyDeviceToDevice, s));
  M->ck_angle = M->angle;
  M->ck_valid = true;
  return FB_OK;
}

int fb_mac_rollback(fb_mac* mac) {
  FB_REQUIRE(mac, FB_ERR_ARG, "NULL handle");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  M->last_stats_valid = false;
  FB_REQUIRE(M->ck_valid, FB_ERR_STATE, "no checkpoint to roll back to");
  FB_HIP(hipSetDevice(M->ctx->device));
  hipStream_t s = M->ctx->stream;
  FB_HIP(hipMemcpyAsync(M->U[M->cur], M->CKU, M->NU * M->sz, hipMemcpyDeviceToDevice, s));
  FB_HIP(hipMemcpyAsync(M->V[M->cur], M->CKV, M->NV * M->sz, hipMemcpyDeviceToDevice, s));
  if (M->geom_valid && M->angle != M->ck_angle) {
    fb_pose pose{};
    pose.angle_degrees = M->ck_angle;
    // restore the checkpointed mask and rebuild the operator from it
    FB_HIP(hipMemcpyAsync(M->solid, M->ck_solid, M->N, hipMemcpyDeviceToDevice, s));
    M->angle = M->ck_angle;
    int rc = M->poisson.setup(s, M->dom.nx, M->dom.ny, M->g.dx, M->g.dy, M->dom.periodic_x, M->dom.periodic_y,
                              M->solid);
    if (rc) return rc;
    // the signed distance is recomputed lazily at the next pose change
  }
  FB_HIP(hipStreamSynchronize(s));
  return FB_OK;
}

int fb_sf_substep(fb_mac* mac, const fb_sf_params* params, const fb_pose* pose, fb_sf_report* report) {
  FB_REQUIRE(mac && params && pose && report, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_REQUIRE(params->max_iterations >= 1, FB_ERR_ARG, "max_iterations must be positive");
  FB_REQUIRE(params->mixed_poisson >= 0 && params->mixed_poisson <= 2, FB_ERR_ARG, "unknown pressure precision %d",
             (int)params->mixed_poisson);
  FB_REQUIRE(params->warm_pressure >= 0 && params->warm_pressure <= 2 &&
                 !(params->warm_pressure && params->mixed_poisson == 1),
             FB_ERR_ARG, "warm_pressure must be 0, 1 or 2, and 0 with mixed_poisson 1");
  FB_REQUIRE(params->face_transport >= 0 && params->face_transport <= 2, FB_ERR_ARG,
             "face_transport must be 0 (cell), 1 (face advection) or 2 (skew-RK2)");
  FB_HIP(hipSetDevice(M->ctx->device));
  *report = fb_sf_report{};
  return MAC_DISPATCH(M, fb::substep, *M, *params, *pose, report);
}

int fb_mac_project(fb_mac* mac, const fb_sf_params* params, fb_sf_report* report) {
  FB_REQUIRE(mac && params && report, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_REQUIRE(M->geom_valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(M->ctx->device));
  *report = fb_sf_report{};
  return MAC_DISPATCH(M, fb::project, *M, *params, params->dt, report);
}

int fb_mac_get_stats(fb_mac* mac, double wall_omega_rad, fb_mac_stats* stats) {
  FB_REQUIRE(mac && stats, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_REQUIRE(M->geom_valid, FB_ERR_STATE, "no geometry has been set");
  if (M->last_stats_valid && M->last_stats_omega == wall_omega_rad) {
    // the faces have not changed since the projection that measured them
    *stats = M->last_stats;
    return FB_OK;
  }
  FB_HIP(hipSetDevice(M->ctx->device));
  return MAC_DISPATCH(M, fb::mac_stats, *M, wall_omega_rad, stats);
}

// The exported cell velocity of the grid: face averages in the face dtype (fb/solvers/stable_fluids.py
// cell_velocity), with the grid's mask.
static fb::DiagField mac_diag_field(const fb::MacGrid& M) {
  fb::DiagField f;
  f.nx = M.g.nx;
  f.ny = M.g.ny;
  f.px = M.g.px;
  f.py = M.g.py;
  f.x0 = M.g.x0;
  f.y0 = M.g.y0;
  f.dx = M.g.dx;
  f.dy = M.g.dy;
  f.kind = fb::kDiagFaces;
  f.storage_f64 = f.arith_f64 = M.sz == sizeof(double);
  f.u = M.U[M.cur];
  f.v = M.V[M.cur];
  f.solid = M.solid;
  return f;
}

int fb_mac_diagnostics(fb_mac* mac, const fb_diag_params* params, fb_diag* out) {
  FB_REQUIRE(mac && params && out, FB_ERR_ARG, "NULL argument");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_REQUIRE(M->geom_valid, FB_ERR_STATE, "no geometry has been set");
  FB_HIP(hipSetDevice(M->ctx->device));
  *out = fb_diag{};
  return fb::field_diagnostics(M->ctx, mac_diag_field(*M), *params, out);
}

int fb_mac_sample_velocity(fb_mac* mac, int64_t n, const double* points, void* out) {
  FB_REQUIRE(mac, FB_ERR_ARG, "NULL handle");
  auto* M = reinterpret_cast<MacGrid*>(mac);
  FB_HIP(hipSetDevice(M->ctx->device));
  return fb::sample_cell_velocity(M->ctx, mac_diag_field(*M), n, points, out);
}

}  // extern "C"
