// Device diagnostics shared by the three solvers: the run-level field metrics of
// fb/core/metrics.py:24-104 (kinetic energy, enstrophy, cell divergence, wake width, recirculation
// area), the lattice's density mean and cut-link adjacent normal speed (fb/solvers/lbm.py:1054-1146),
// the PIC/FLIP particle census over fluid cells (fb/solvers/pic_flip.py:1115-1136) and the bilinear
// velocity sample of fb/core/interpolation.py:9-79, computed where the fields live.
#pragma once

#include "fb_common.h"

namespace fb {

// The cell velocity a solver exports, read in place: cell values (element stride 1 for planes, 2
// for an interleaved (ny, nx, 2) field) or the cell average of MAC faces u (ny, nx+1), v (ny+1, nx).
enum DiagKind { kDiagCells = 0, kDiagFaces = 1 };

struct DiagField {
  int nx = 0, ny = 0, px = 0, py = 0;
  double x0 = 0.0, y0 = 0.0, dx = 1.0, dy = 1.0;
  int kind = kDiagCells;
  int storage_f64 = 1;  // u, v and rho are double (1) or float (0)
  int arith_f64 = 1;    // dtype of the exported field, in which the per-cell terms are formed
  const void* u = nullptr;
  const void* v = nullptr;
  int stride = 1;
  const uint8_t* solid = nullptr;
  const void* rho = nullptr;      // lattice density (storage type), or null
  const double* sdf = nullptr;    // cell signed distance for the cut-link speed, or null
  double pivot_x = 0.0, pivot_y = 0.0;
};

// Field metrics into *out (host), on the context's stream; returns when they are on the host.
int field_diagnostics(Ctx* c, const DiagField& f, const fb_diag_params& p, fb_diag* out);

// Bilinear sample of the cell velocity at n host points (n, 2) float64 into out_host (n, 2) of the
// exported dtype.
int sample_cell_velocity(Ctx* c, const DiagField& f, int64_t n, const double* points_host, void* out_host);

// Particle census: particles per cell by fb/core/particle_population.py:19-31 (cell index in the
// positions' dtype, clipped to the grid), histogram over fluid cells into hist_host
// (FB_DIAG_HISTOGRAM entries, the last one counting every cell at or above it), the fluid, empty and
// underfilled cell counts and the maximum into *out; the per-cell counts into counts_host when given.
int particle_census(Ctx* c, const DiagField& f, int64_t n, const void* positions, fb_diag* out,
                    int64_t* hist_host, int32_t* counts_host);

}  // namespace fb
