// MAC-grid handle shared by the Stable Fluids substep and the PIC/FLIP grid stages.
#pragma once

#include "geometry.h"
#include "poisson.h"

namespace fb {

#pragma clang fp contract(off)

struct MacGeom {
  int nx, ny, px, py;
  double x0, y0, dx, dy;
  double x1, y1;  // far bounds (the last face coordinates)
  // cell-centre linspace coefficients
  double sx, stx, ex, sy, sty, ey;
};

inline MacGeom make_geom(const fb_domain& d) {
  MacGeom g;
  g.nx = d.nx;
  g.ny = d.ny;
  g.px = d.periodic_x;
  g.py = d.periodic_y;
  g.x0 = d.x0;
  g.y0 = d.y0;
  g.dx = (d.x1 - d.x0) / d.nx;
  g.dy = (d.y1 - d.y0) / d.ny;
  g.x1 = d.x1;
  g.y1 = d.y1;
  CellGrid c = make_cell_grid(d);
  g.sx = c.sx;
  g.stx = c.stx;
  g.ex = c.ex;
  g.sy = c.sy;
  g.sty = c.sty;
  g.ey = c.ey;
  return g;
}

#pragma clang fp contract(fast)

struct MacGrid;

// Replacement for the grid's own pressure solve (the request loop batches solves of several grids):
// solves the grid's rhs into its pres, with the report complete on return.
struct SolveHook {
  int (*fn)(void* user, MacGrid& M, double tolerance, int max_iterations, int mode, PoissonReport* rep) = nullptr;
  void* user = nullptr;
};

struct MacGrid {
  Ctx* ctx = nullptr;
  SolveHook hook{};
  fb_domain dom{};
  fb_foil foil{};
  MacGeom g{};
  int prec = FB_F64;
  size_t sz = 8;
  size_t N = 0, NU = 0, NV = 0;
  void *U[2] = {}, *V[2] = {};  // current faces and scratch
  int cur = 0;
  void *CU[4] = {}, *CV[4] = {};  // cell planes: velocity, first, corrected, spare
  void *DU[3] = {}, *DV[3] = {};  // diffusion: orig, ping, pong (face-sized to serve PIC)
  void *CKU = nullptr, *CKV = nullptr;  // checkpoint faces
  double* sdf = nullptr;
  uint8_t* solid = nullptr;
  uint8_t* ck_solid = nullptr;
  double angle = 0.0, ck_angle = 0.0;
  bool geom_valid = false, ck_valid = false;
  double* rhs = nullptr;
  double* pres = nullptr;
  double* pres_prev = nullptr;  // extrapolated warm starts: the pressure of the solve before the last
  unsigned long long* skew_key = nullptr;  // skew-RK2 advection rate (order-preserving key)
  int pres_solves = 0;          // solves whose results sit in pres and pres_prev since the last reset
  unsigned long long* keys = nullptr;
  unsigned long long* host_keys = nullptr;
  unsigned long long* resid = nullptr;  // 2 x (maxit + 2)
  unsigned long long* host_resid = nullptr;
  int resid_cap = 0;
  int last_sweeps = 0;  // Jacobi sweeps of the previous diffusion (sets the first polling chunk)
  Poisson poisson;
  bool poisson_ready = false;
  // statistics of the faces after the last projection (valid until the faces change)
  fb_mac_stats last_stats{};
  double last_stats_omega = 0.0;
  bool last_stats_valid = false;
};

int create_mac(Ctx* c, const fb_domain* dom, const fb_foil* foil, int precision, MacGrid** out);
void free_mac(MacGrid* M);
int mac_geometry(MacGrid& M, const fb_pose& pose);
template <class T>
int diffuse(MacGrid& M, const T* in0, const T* in1, int W0, int H0, int P0, int W1, int H1, int P1, double nu,
            double dt, double tol, int maxit, T** result, fb_sf_report* rep);
template <class T>
int project(MacGrid& M, const fb_sf_params& p, double dt, fb_sf_report* rep);
template <class T>
int mac_stats(MacGrid& M, double w, fb_mac_stats* st);
template <class T>
int launch_faces_to_cell(MacGrid& M, const T* u, const T* v, T* cu, T* cv);
template <class T>
int launch_cell_to_faces(MacGrid& M, const T* cu, const T* cv, T* u, T* v);
template <class T>
int launch_dup_copy(MacGrid& M, T* f, int rows, int cols, int pitch, int dup_x, int dup_y);
// One Stable Fluids substep (the body of fb_sf_substep) on M's stream.
int mac_substep(MacGrid& M, const fb_sf_params& p, const fb_pose& pose, fb_sf_report* rep);
// Moves M onto context c: its kernels, copies and reduction slots then use c's stream. The
// caller makes sure no work of M is pending on the previous stream.
void mac_bind(MacGrid& M, Ctx* c);
// Whether the pressure solve replays graphs of its CG iteration chunks.
void mac_use_graphs(MacGrid& M, bool on);

}  // namespace fb
