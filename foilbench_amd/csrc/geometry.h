// NACA four-digit signed distance on the device, evaluated in the reference's FP64
// operation order (fb/core/geometry.py:35-44, 60-106) with FP contraction off.
//
// x**3 and x**4 go through libm pow in the reference, whose last bit is platform
// dependent; here they are correctly rounded (double-double products). The sign of the
// distance, and so every mask, agrees with the reference except for points within about
// one ulp of the surface.
#pragma once

#include "fb_common.h"

namespace fb {

struct NacaDev {
  double chord, pivot_x, pivot_y, half_scale, camber_fwd, camber_aft, two_p, one_minus_two_p, p;
  int cambered;
  double cos_a, sin_a;
};

inline NacaDev make_naca(const fb_foil& f, const fb_pose& pose) {
  NacaDev g;
  g.chord = f.chord;
  g.pivot_x = f.pivot_x;
  g.pivot_y = f.pivot_y;
  g.half_scale = f.half_scale;
  g.camber_fwd = f.camber_fwd;
  g.camber_aft = f.camber_aft;
  g.two_p = f.two_p;
  g.one_minus_two_p = f.one_minus_two_p;
  g.p = f.p;
  g.cambered = f.cambered;
  g.cos_a = pose.cos_a;
  g.sin_a = pose.sin_a;
  return g;
}

#pragma clang fp contract(off)

// Correctly rounded x^3 and x^4 for x in [0, 1] from error-free products.
__device__ __forceinline__ double cr_cube(double x) {
  double h = x * x;
  double l = __fma_rn(x, x, -h);
  double p = h * x;
  double e = __fma_rn(h, x, -p);
  return p + (e + l * x);
}
__device__ __forceinline__ double cr_quart(double x) {
  double h = x * x;
  double l = __fma_rn(x, x, -h);
  double p = h * h;
  double e = __fma_rn(h, h, -p);
  return p + (e + 2.0 * h * l);
}

__device__ __forceinline__ double naca_sdf(const NacaDev& g, double px, double py) {
  // _to_local (geometry.py:35-44)
  double tx = px - g.pivot_x;
  double ty = py - g.pivot_y;
  double lx = g.cos_a * tx + g.sin_a * ty;
  double ly = (-g.sin_a) * tx + g.cos_a * ty;
  lx = lx + 0.25 * g.chord;
  // surfaces (geometry.py:60-90)
  double x = fmin(fmax(lx / g.chord, 0.0), 1.0);
  double x2 = x * x;
  double poly = 0.2969 * sqrt(fmax(x, 0.0)) - 0.1260 * x - 0.3516 * x2 + 0.2843 * cr_cube(x) -
                0.1036 * cr_quart(x);
  double half = g.half_scale * poly;
  double camber = 0.0;
  if (g.cambered) {
    if (x < g.p)
      camber = g.camber_fwd * (g.two_p * x - x2);
    else
      camber = g.camber_aft * ((g.one_minus_two_p + g.two_p * x) - x2);
    camber = camber * g.chord;
  }
  double upper = camber + half;
  double lower = camber - half;
  // signed_distance (geometry.py:92-106)
  double y = ly;
  bool inside_x = (lx >= 0.0) && (lx <= g.chord);
  double outside_v = fmax(y - upper, lower - y);
  double inside_v = -fmin(upper - y, y - lower);
  double vertical = ((y <= upper) && (y >= lower)) ? inside_v : outside_v;
  double ddx = fmax(-lx, lx - g.chord);
  double outside_x = fmax(ddx, 0.0);
  return inside_x ? vertical : hypot(outside_x, fmax(vertical, 0.0));
}

// np.linspace(start, stop, n)[i] (numpy/_core/function_base.py: i*step + start, last = stop).
__device__ __forceinline__ double linspace_at(int i, int n, double start, double step, double stop) {
  return (i == n - 1) ? stop : ((double)i * step + start);
}

#pragma clang fp contract(fast)


#pragma clang fp contract(off)

// Signed distance of particle positions in the particle dtype T, with the reference's NumPy
// promotions (fb/core/geometry.py:35-106 on a (P, 2) array of dtype T): the rotation uses
// FP64 cos/sin scalars and rounds to T, everything after it is T arithmetic with Python-float
// coefficients rounded to T. For T = double this is naca_sdf.
template <class T>
__device__ __forceinline__ T naca_sdf_t(const NacaDev& g, T px, T py) {
  if constexpr (sizeof(T) == 8) {
    return (T)naca_sdf(g, (double)px, (double)py);
  } else {
    T tx = px - (T)g.pivot_x;
    T ty = py - (T)g.pivot_y;
    T lx = (T)(g.cos_a * (double)tx + g.sin_a * (double)ty);
    T ly = (T)((-g.sin_a) * (double)tx + g.cos_a * (double)ty);
    lx = lx + (T)(0.25 * g.chord);
    const T chord = (T)g.chord;
    T x = fminf(fmaxf(lx / chord, (T)0), (T)1);
    T x2 = x * x;
    T x3 = (T)((double)x * (double)x * (double)x);
    T x4 = (T)((double)x2 * (double)x2);
    T poly = (T)0.2969 * sqrtf(fmaxf(x, (T)0)) - (T)0.1260 * x - (T)0.3516 * x2 + (T)0.2843 * x3 - (T)0.1036 * x4;
    T half = (T)g.half_scale * poly;
    T camber = (T)0;
    if (g.cambered) {
      if (x < (T)g.p)
        camber = (T)g.camber_fwd * ((T)g.two_p * x - x2);
      else
        camber = (T)g.camber_aft * (((T)g.one_minus_two_p + (T)g.two_p * x) - x2);
      camber = camber * chord;
    }
    T upper = camber + half;
    T lower = camber - half;
    T y = ly;
    bool inside_x = (lx >= (T)0) && (lx <= chord);
    T outside_v = fmaxf(y - upper, lower - y);
    T inside_v = -fminf(upper - y, y - lower);
    T vertical = ((y <= upper) && (y >= lower)) ? inside_v : outside_v;
    T ddx = fmaxf(-lx, lx - chord);
    T outside_x = fmaxf(ddx, (T)0);
    return inside_x ? vertical : hypotf(outside_x, fmaxf(vertical, (T)0));
  }
}

// normals (fb/core/geometry.py:108-126) in the particle dtype.
template <class T>
__device__ __forceinline__ void naca_normal_t(const NacaDev& g, T px, T py, T& nx_out, T& ny_out) {
  const double epsd = fmax(g.chord * 1.0e-4, 1.0e-6);
  const T eps = (T)epsd;
  T gx = naca_sdf_t<T>(g, px + eps, py + (T)0) - naca_sdf_t<T>(g, px - eps, py - (T)0);
  T gy = naca_sdf_t<T>(g, px + (T)0, py + eps) - naca_sdf_t<T>(g, px - (T)0, py - eps);
  T len = sqrt(gx * gx + gy * gy);
  if (len < eps) {
    gx = (T)(-g.sin_a);
    gy = (T)g.cos_a;
    len = (T)1;
  }
  T den = len > eps ? len : eps;
  nx_out = gx / den;
  ny_out = gy / den;
}

#pragma clang fp contract(fast)

struct CellGrid {
  int nx, ny;
  double sx, stx, ex;  // x centres: start, step, stop
  double sy, sty, ey;
};

inline CellGrid make_cell_grid(const fb_domain& d) {
  CellGrid c;
  c.nx = d.nx;
  c.ny = d.ny;
  double dx = (d.x1 - d.x0) / d.nx;
  double dy = (d.y1 - d.y0) / d.ny;
  c.sx = d.x0 + 0.5 * dx;
  c.ex = d.x1 - 0.5 * dx;
  c.stx = (c.ex - c.sx) / (d.nx - 1);
  c.sy = d.y0 + 0.5 * dy;
  c.ey = d.y1 - 0.5 * dy;
  c.sty = (c.ey - c.sy) / (d.ny - 1);
  return c;
}

// Launch: sdf (ny*nx) and solid bytes for cell centres.
int launch_cell_sdf(hipStream_t s, const CellGrid& cg, const NacaDev& g, double* sdf, uint8_t* solid);

}  // namespace fb
