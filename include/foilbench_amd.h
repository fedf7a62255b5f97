/*
 * foilbench_amd: C ABI of the MI355X (gfx950) per-step grid update of FoilBench's three
 * 2D solvers. Plain pointers, sizes and POD structs only; every entry point returns an
 * int status (FB_OK = 0) and fb_last_error() describes the last failure of the calling
 * thread.
 *
 * The reference (Gravifer/FoilBench, Python) has no FFI: its hot path is Python calling
 * numba kernels and NumPy/SciPy stages in-process. Each entry point below names the
 * reference function(s) it replaces (paths relative to implementations/python/src/;
 * fb/ = foilbench_py/, kern/ = foilbench_kernels/). The Python binding a maintainer adds on
 * the reference side is the ctypes layer in foilbench_amd/_native.py.
 *
 * State lives on the device for the lifetime of a handle; host arrays passed in are
 * copied and never retained. Host arrays use the reference's C-order layouts:
 * populations (ny, nx, 9), faces u (ny, nx+1) and v (ny+1, nx), fields (ny, nx[, 2]).
 */
#ifndef FOILBENCH_AMD_H
#define FOILBENCH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fb_status {
  FB_OK = 0,
  FB_ERR_ARG = 1,      /* invalid argument (ValueError/TypeError in the reference) */
  FB_ERR_HIP = 2,      /* HIP runtime failure */
  FB_ERR_STATE = 3,    /* handle used out of order (RuntimeError) */
  FB_ERR_NUMERIC = 4,  /* classified numerical failure; see the report struct */
  FB_ERR_NOMEM = 5
};

/* Storage/arithmetic forms.
 * F64: FP64 storage and arithmetic (the plain drop-in form).
 * F32: FP32 storage with FP64 kernel arithmetic, like the reference's float32 scenarios.
 * F32_MIXED / F16_MIXED: populations stored shifted by their lattice weights (f - w) in
 *   FP32 / FP16, collided in FP64 / FP32 arithmetic (the precision forms). */
enum fb_precision { FB_F64 = 0, FB_F32 = 1, FB_F32_MIXED = 2, FB_F16_MIXED = 3 };

typedef struct fb_ctx fb_ctx;
typedef struct fb_lbm fb_lbm;

/* Uniform cell grid: cells span [x0,x1] x [y0,y1] (fb/core/models.py:50-93). */
typedef struct fb_domain {
  int32_t nx, ny;
  int32_t periodic_x, periodic_y;
  double x0, x1, y0, y1;
} fb_domain;

/* NACA four-digit foil with the scalar coefficients of fb/core/geometry.py:60-90
 * evaluated by the host exactly as the reference does in Python floats. */
typedef struct fb_foil {
  double chord, pivot_x, pivot_y;
  double half_scale;      /* 5.0 * thickness * chord */
  double camber_fwd;      /* m / p**2 */
  double camber_aft;      /* m / (1.0 - p)**2 */
  double two_p;           /* 2.0 * p */
  double one_minus_two_p; /* 1.0 - 2.0 * p */
  double p;               /* camber position */
  int32_t cambered;       /* m > 0 and p > 0 */
  int32_t pad_;
} fb_foil;

/* Rigid pose: cos and sin of np.deg2rad(angle_degrees), evaluated by the host with NumPy
 * (fb/core/geometry.py:35-44). */
typedef struct fb_pose {
  double angle_degrees, cos_a, sin_a;
} fb_pose;

const char* fb_last_error(void);
int fb_version(void);
int fb_device_count(int* count);

int fb_ctx_create(int device, fb_ctx** out);
int fb_ctx_destroy(fb_ctx* ctx);
int fb_ctx_synchronize(fb_ctx* ctx);
/* Device-to-device streaming copy of `bytes` (reps times), timed with HIP events; used to
 * calibrate the HBM byte counters and the achievable bandwidth. No reference counterpart. */
int fb_copy_benchmark(fb_ctx* ctx, int64_t bytes, int reps, double* ms_per_copy);

/* Signed distance and mask of cell centres (fb/core/geometry.py:92-106, 157-168) into
 * host arrays sdf (ny, nx) float64 and solid (ny, nx) uint8. */
int fb_naca_cell_sdf(fb_ctx* ctx, const fb_domain* dom, const fb_foil* foil, const fb_pose* pose,
                     double* sdf_host, uint8_t* solid_host);

/* Signed distance and normals of arbitrary points (fb/core/geometry.py:92-126); points are
 * (n, 2) float64; normals may be NULL. */
int fb_naca_points(fb_ctx* ctx, const fb_foil* foil, const fb_pose* pose, int64_t n,
                   const double* points_host, double* sdf_host, double* normals_host);

/* ------------------------------------------------------------------ D2Q9 TRT LBM ---- */
/* Replaces kern/lbm.py:12-168 (trt_collision_kernel, moving_wall_stream_kernel,
 * sponge_kernel) and the boundary/step logic of fb/solvers/lbm.py:486-649, 463-484,
 * 357-368. One handle = one lattice with its geometry and transactional checkpoint. */

typedef struct fb_lbm_params {
  double omega_plus, omega_minus; /* TRT rates (fb/solvers/lbm.py:311-338) */
  double u_lattice;               /* convective outlet speed (lbm.py:637-639) */
  double target_ux, target_uy;    /* Zou-He target, storage-rounded (lbm.py:599-600) */
  double wall_omega_rad;          /* foil angular velocity in rad/s (lbm.py:587) */
  double wall_scale;              /* u_lattice / reference speed (lbm.py:606) */
  double velocity_scale;          /* reference speed / u_lattice (lbm.py:457-461) */
  double feq_boundary[9];         /* f_eq(1, target) in storage dtype (lbm.py:340-355) */
  double feq_uncovered[9];        /* f_eq of newly uncovered cells (lbm.py:476-481) */
} fb_lbm_params;

typedef struct fb_lbm_stats {
  double min_population;     /* min over all populations (lbm.py:736) */
  double max_speed;          /* max physical |u| over all cells (lbm.py:777) */
  double density_excursion;  /* max |rho-1| over fluid cells (lbm.py:760) */
  int64_t nonfinite_populations;
  int64_t bad_fluid_density; /* non-finite or <= 0 fluid densities (lbm.py:754) */
  int64_t nonfinite_velocity;
} fb_lbm_stats;

/* sponge_y (ny) and sponge_x (nx): float64 strength profiles; the cell strength is
 * max(sponge_y[y], sponge_x[x]) (fb/solvers/lbm.py:270-294). */
int fb_lbm_create(fb_ctx* ctx, const fb_domain* dom, const fb_foil* foil, int precision,
                  int channel_walls, const double* sponge_y, const double* sponge_x, fb_lbm** out);
int fb_lbm_destroy(fb_lbm* lbm);
/* Pre-collision populations (ny, nx, 9) and outlet history (ny, 9), host dtype double for
 * FB_F64 and float otherwise (mixed forms take unshifted values). */
int fb_lbm_set_state(fb_lbm* lbm, const void* f_host, const void* outlet_host);
int fb_lbm_get_state(fb_lbm* lbm, void* f_host, void* outlet_host);
/* Pre-collision state from density (ny, nx) and lattice velocity (ny, nx, 2), host dtype:
 * f = f_eq(rho, u) evaluated on the device (fb/solvers/lbm.py:252-254, 428-445, 994-995);
 * the outlet history becomes f[:, -1] (lbm.py:254). */
int fb_lbm_set_equilibrium(fb_lbm* lbm, const void* rho_host, const void* vel_host);
/* Geometry for a pose without resetting uncovered cells (initialize/import). */
int fb_lbm_set_pose(fb_lbm* lbm, const fb_pose* pose);
/* Pose change outside an advance (fb/solvers/lbm.py:463-484, used by restart): new geometry,
 * and cells uncovered by the move take params->feq_uncovered. No-op for the same angle. */
int fb_lbm_update_pose(fb_lbm* lbm, const fb_pose* pose, const fb_lbm_params* params);
int fb_lbm_get_geometry(fb_lbm* lbm, double* sdf_host, uint8_t* solid_host);
/* Transaction: the state at checkpoint is kept intact until the next checkpoint. */
int fb_lbm_checkpoint(fb_lbm* lbm);
int fb_lbm_rollback(fb_lbm* lbm);
/* f <- f_eq(rho, u*ratio) + ratio*(f - f_eq(rho, u)); outlet <- f[:, -1]
 * (fb/solvers/lbm.py:357-368). */
int fb_lbm_rescale(fb_lbm* lbm, double ratio);
/* Run `substeps` lattice steps; poses[k] is the foil pose of substep k
 * (fb/solvers/lbm.py:714-722). The stats of the final state are written to *stats. */
int fb_lbm_advance(fb_lbm* lbm, const fb_lbm_params* params, int substeps, const fb_pose* poses,
                   fb_lbm_stats* stats);
/* Postcondition reductions of the current state (fb/solvers/lbm.py:687-688, 736-793). */
int fb_lbm_stats_now(fb_lbm* lbm, const fb_lbm_params* params, fb_lbm_stats* stats);
/* Density (ny, nx) and physical velocity (ny, nx, 2) in the host dtype
 * (fb/solvers/lbm.py:447-461). */
int fb_lbm_macroscopic(fb_lbm* lbm, double velocity_scale, void* rho_host, void* vel_host);
/* Force on the foil per unit depth in lattice units, by momentum exchange on the pre-collision
 * state: the sum over fluid cells x and directions e with a solid neighbour x + c_e of
 * c_e (f_e(x) + f_opp(e)(x)). Build-defined: the reference computes no lift or drag.
 * force[0], force[1]: x and y components. Coefficients: C = 2 F / (u_lattice^2 * chord cells). */
int fb_lbm_forces(fb_lbm* lbm, double* force);
/* Timing harness: run `steps` lattice steps with fixed parameters and no pose change,
 * without postcondition reductions; returns once queued (use fb_ctx_synchronize). */
int fb_lbm_run_fixed(fb_lbm* lbm, const fb_lbm_params* params, int steps);
/* Same, timed with one HIP event pair on the handle's stream around the `steps` back-to-back
 * step launches; elapsed_ms / steps is the average launch duration (gaps included). */
int fb_lbm_time_fixed(fb_lbm* lbm, const fb_lbm_params* params, int steps, double* elapsed_ms);

/* ------------------------------------------------------------ Stable Fluids (MAC) ---- */
/* Replaces the MAC-grid stages of fb/solvers/stable_fluids.py:424-494 and fb/core/grid.py
 * (faces_to_cell, advect_velocity/rk2_backtrace/local_velocity_bounds, implicit_diffuse,
 * cell_to_faces, apply_domain_boundaries, enforce_solid_faces, project_faces) with the
 * pressure solve of fb/core/_scipy_adapter.py:16-161. Faces: u (ny, nx+1), v (ny+1, nx)
 * in the storage dtype (FB_F64 or FB_F32); the pressure solve is FP64. */

typedef struct fb_mac fb_mac;

/* Substep outcome (the reference's classified failures of one substep). */
enum fb_sf_status {
  FB_SF_OK = 0,
  FB_SF_VISCOSITY_NOT_CONVERGED = 1,     /* convergence_failure, stage viscosity */
  FB_SF_NONFINITE_PROJECTION_INPUT = 2,  /* nonfinite_state, stage projection */
  FB_SF_PROJECTION_CFL = 3,              /* excessive_velocity, stage projection */
  FB_SF_NONFINITE_RHS = 4,               /* nonfinite_state (pressure RHS) */
  FB_SF_RHS_RANGE = 5,                   /* projection_failure (RHS above safe range) */
  FB_SF_NONFINITE_ITERATE = 6,           /* nonfinite_state (CG iterate) */
  FB_SF_CG_BREAKDOWN = 7,                /* projection_failure (CG floating-point failure) */
  FB_SF_PROJECTION_NOT_CONVERGED = 8,    /* projection_failure (not converged) */
  FB_SF_NONFINITE_PROJECTION_OUTPUT = 9  /* nonfinite_state after projection */
};

typedef struct fb_sf_params {
  double dt;               /* substep (or projection pseudo-step) */
  double viscosity;        /* U_ref * chord / Re */
  double tolerance;        /* pressure_tolerance (also the Jacobi tolerance) */
  int32_t max_iterations;  /* pressure_max_iterations (also the Jacobi cap) */
  int32_t maccormack;      /* 1: limited MacCormack, 0: semi-Lagrangian */
  double freestream_x, freestream_y;
  double wall_omega_rad;   /* foil angular velocity in rad/s */
  double cfl_limit;        /* projection CFL limit max(1, 2*stable_cfl) */
  int32_t channel_walls;   /* Poiseuille channel walls */
  int32_t mixed_poisson;   /* 0: FP64 PCG; 1: FP32 CG sweeps inside FP64 residual refinement;
                              2: FP64 PCG with its DST-I preconditioner in FP32 */
  int32_t warm_pressure;   /* 0: CG starts from zero (the reference); 1: from the grid's previous
                              pressure p_n; 2: from 2 p_n - p_(n-1) once two solves are on record
                              (SciPy cg with x0: r0 = b - A x0, same stopping rule); a projection
                              that fails zeroes the pressure; not with mixed_poisson 1 */
  int32_t face_transport;  /* 0: cell-centred advection (maccormack selects the limiter);
                              1: face-native advection, stable_face_advection (fb/core/grid.py:350-376);
                              2: skew-symmetric RK2, stable_advection = "skew-rk2" (grid.py:481-503);
                              1 and 2 diffuse on the faces (grid.py:278-303) */
} fb_sf_params;

typedef struct fb_sf_report {
  int32_t status;  /* enum fb_sf_status */
  int32_t viscosity_iterations;
  double viscosity_residual;
  int32_t viscosity_converged;
  int32_t pressure_iterations;
  double pressure_residual;
  int32_t pressure_info;
  int32_t pressure_converged;
  int32_t pressure_refinements;
  int32_t pad_;
  double projection_cfl;
} fb_sf_report;

typedef struct fb_mac_stats {
  double max_face_speed;   /* max(|u|, |v|) (stable_fluids.py:276) */
  double max_wall_speed;   /* max wall speed over solid cells (:277-279) */
  double max_cell_speed;   /* max |faces_to_cell| (:337, :502) */
  double divergence_linf;  /* fluid MAC divergence (fb/core/grid.py:100-111) */
  double solid_leakage;    /* fluid-solid face leakage (fb/core/grid.py:114-131) */
  int64_t nonfinite_faces;
  int64_t nonfinite_wall;
} fb_mac_stats;

int fb_mac_create(fb_ctx* ctx, const fb_domain* dom, const fb_foil* foil, int precision, fb_mac** out);
int fb_mac_destroy(fb_mac* mac);
int fb_mac_set_faces(fb_mac* mac, const void* u_host, const void* v_host);
int fb_mac_get_faces(fb_mac* mac, void* u_host, void* v_host);
/* Mask for a pose (fb/core/geometry.py:157-159); rebuilds the masked Poisson operator. */
int fb_mac_set_pose(fb_mac* mac, const fb_pose* pose);
int fb_mac_get_geometry(fb_mac* mac, double* sdf_host, uint8_t* solid_host);
int fb_mac_get_pressure(fb_mac* mac, double* pressure_host);
int fb_mac_checkpoint(fb_mac* mac);
int fb_mac_rollback(fb_mac* mac);
/* One Stable Fluids substep (fb/solvers/stable_fluids.py:468-494): advect, implicit viscosity,
 * cell_to_faces, mask at `pose`, projection with the CFL guard (:256-330). */
int fb_sf_substep(fb_mac* mac, const fb_sf_params* params, const fb_pose* pose, fb_sf_report* report);
/* Projection only (fb/solvers/stable_fluids.py:256-330; initialize and import). */
int fb_mac_project(fb_mac* mac, const fb_sf_params* params, fb_sf_report* report);
int fb_mac_get_stats(fb_mac* mac, double wall_omega_rad, fb_mac_stats* stats);
/* Skew-RK2 advection rate of the current faces (fb/core/grid.py:434-443): the largest
 * |u|/dx + |v on u|/dy over u faces and |u on v|/dx + |v|/dy over v faces, in the storage dtype. */
int fb_mac_skew_rate(fb_mac* mac, double* rate);
/* Pressure (ny, nx) FP64 of the last projection, replaced (import, tests). */
int fb_mac_set_pressure(fb_mac* mac, const double* pressure_host);
/* Force on the foil per unit depth at unit density, in FP64 (build-defined: the reference computes
 * no lift or drag). On every fluid-solid cell interface the fluid cell's pressure acts along the
 * interface normal, and viscosity * (tangential fluid velocity at the cell centre - wall velocity)
 * / (half a cell) acts along it. force[0], force[1]: x and y components. */
int fb_mac_forces(fb_mac* mac, double viscosity, double wall_omega_rad, double* force);

/* ---------------------------------------------------------------------- PIC/FLIP ---- */
/* Replaces the particle half of fb/solvers/pic_flip.py (seeding :317-333, collisions
 * :527-618, advection :481-525, the FLIP update :765-840) and the quadratic B-spline
 * transfers kern/pic.py:105-205 (via fb/solvers/_numba_adapter.py:108-160). The grid half
 * (implicit viscosity on faces, projection) runs on the handle's MAC grid. Particles are
 * (P, 2) arrays in the storage dtype; the grid velocity is two (ny, nx) planes. The PCG32
 * stream (fb/core/rng.py) lives on the device with the particles. */

typedef struct fb_pic fb_pic;

typedef struct fb_pic_params {
  double dt;               /* substep */
  double viscosity;        /* U_ref * chord / Re */
  double tolerance;        /* pressure_tolerance (also the Jacobi tolerance) */
  int32_t max_iterations;  /* pressure_max_iterations */
  int32_t channel_walls;   /* Poiseuille channel walls */
  double freestream_x, freestream_y;
  double wall_omega_rad;   /* boundary angular velocity of the substep control (rad/s) */
  double blend;            /* FLIP fraction of this substep (0 while settling) */
  double wall_travel;      /* |deg2rad(angle - start angle)| * maximum_radius */
  double maximum_radius;   /* NacaFoil.maximum_radius (fb/core/geometry.py:28-33) */
  int32_t mixed_poisson;   /* 0: FP64 PCG; 1: FP32 CG sweeps inside FP64 residual refinement;
                              2: FP64 PCG with its DST-I preconditioner in FP32 */
  int32_t pad_;
} fb_pic_params;

typedef struct fb_pic_report {
  fb_sf_report grid;          /* implicit viscosity and projection of the substep */
  int64_t unsupported_faces;  /* u and v faces whose P2G weight is <= 1e-12 */
  int64_t swept_hits;         /* particles resolved by the swept-collision test */
  int64_t escaped;            /* particles respawned after leaving the domain */
} fb_pic_report;

typedef struct fb_pic_stats {
  double max_particle_speed;  /* max |particle velocity| (pic_flip.py:701-706) */
  double max_grid_speed;      /* max |grid velocity| */
  int64_t particles_inside;   /* contains() at the pose (pic_flip.py:864-873) */
  int64_t nonfinite;          /* non-finite positions, velocities or grid values */
  int64_t fluid_cells;
  int64_t empty_cells;        /* fluid cells without particles (pic_flip.py:954-961) */
  int64_t underfilled_cells;  /* fluid cells with fewer than 4 particles */
  int64_t particle_count;
} fb_pic_stats;

int fb_pic_create(fb_ctx* ctx, const fb_domain* dom, const fb_foil* foil, int precision, fb_pic** out);
int fb_pic_destroy(fb_pic* pic);
/* The handle's MAC grid (faces of the last projection, mask, pressure) for the fb_mac_* getters. */
int fb_pic_mac(fb_pic* pic, fb_mac** out);
int fb_pic_set_pose(fb_pic* pic, const fb_pose* pose);
int fb_pic_set_grid(fb_pic* pic, const void* gu_host, const void* gv_host);
int fb_pic_get_grid(fb_pic* pic, void* gu_host, void* gv_host);
/* PCG32 stream words (fb/core/rng.py:33-44: checkpoint/restore). */
int fb_pic_set_rng(fb_pic* pic, uint64_t state, uint64_t increment);
int fb_pic_get_rng(fb_pic* pic, uint64_t* state, uint64_t* increment);
/* _seed_particles (fb/solvers/pic_flip.py:317-333): per_cell particles per fluid cell of the
 * current mask, jittered from the device stream; velocities gathered from the grid. */
int fb_pic_seed(fb_pic* pic, int per_cell, int64_t* count);
int fb_pic_count(fb_pic* pic, int64_t* count);
int fb_pic_set_particles(fb_pic* pic, int64_t n, const void* pos_host, const void* vel_host);
int fb_pic_get_particles(fb_pic* pic, void* pos_host, void* vel_host);
/* _particle_to_grid (fb/solvers/pic_flip.py:361-377): P2G with the grid as fallback; the
 * unsupported face fraction; store_grid = 1 replaces the grid with faces_to_cell of the result. */
int fb_pic_particle_to_grid(fb_pic* pic, int store_grid, double* unsupported_fraction);
/* velocities[indices] = G2P(cell_to_faces(grid)) at their positions (pic_flip.py:620-635). */
int fb_pic_regrid(fb_pic* pic, const int64_t* indices, int64_t count);
/* _resolve_particle_collisions (fb/solvers/pic_flip.py:604-618). */
int fb_pic_resolve_collisions(fb_pic* pic, const fb_pose* pose, double wall_omega_rad);
/* One PIC/FLIP substep (fb/solvers/pic_flip.py:765-840): mask at `pose`, collisions, P2G,
 * implicit viscosity, projection, grid update, G2P blend, advection with swept collisions,
 * wrap and respawn, collisions. sweep_table holds (cos, sin) of the swept-collision sample
 * angles start + s/count * (angle - start) for count = 2..16 and s = 1..count (135 pairs). */
int fb_pic_substep(fb_pic* pic, const fb_pic_params* params, const fb_pose* pose, const double* sweep_table,
                   fb_pic_report* report);
int fb_pic_get_stats(fb_pic* pic, const fb_pose* pose, fb_pic_stats* stats);
int fb_pic_checkpoint(fb_pic* pic);
int fb_pic_rollback(fb_pic* pic);

/* particle_to_faces (fb/solvers/_numba_adapter.py:108-135, kern/pic.py:105-161) on host
 * arrays: positions/velocities (n, 2), fallback and output faces u (ny, nx+1), v (ny+1, nx). */
int fb_pic_particle_to_faces(fb_ctx* ctx, int precision, int64_t n, const void* positions, const void* velocities,
                             const void* fallback_u, const void* fallback_v, int nx, int ny, double x0, double y0,
                             double dx, double dy, int periodic_x, int periodic_y, void* u_out, void* v_out,
                             double* unsupported_fraction);
/* fb_pic_particle_to_faces without any allocation, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * Every pointer is a DEVICE pointer: positions/velocities (n, 2), fallback and output faces u (ny, nx+1),
 * v (ny+1, nx), and unsupported_faces, one uint64 that receives the count of u and v faces whose P2G weight
 * is <= 1e-12 (the fraction is that count over ny*(nx+1) + (ny+1)*nx). The work is queued on the context's
 * stream and the call returns without synchronising (fb_ctx_synchronize waits for it). The counting-sort
 * buffers live in a fixed device array of the library (about 23 MiB per GPU, held until the process ends):
 * at most 524,288 particles and 262,144 cells; larger inputs are rejected with FB_ERR_ARG. All calls on one
 * GPU share that array, so calls that overlap on one GPU race: serialise them (one stream, or wait between
 * calls). Same kernels and results as fb_pic_particle_to_faces. */
int fb_pic_particle_to_faces_noalloc(fb_ctx* ctx, int precision, int64_t n, const void* positions,
                                     const void* velocities, const void* fallback_u, const void* fallback_v, int nx,
                                     int ny, double x0, double y0, double dx, double dy, int periodic_x,
                                     int periodic_y, void* u_out, void* v_out, uint64_t* unsupported_faces);
/* faces_to_particle (fb/solvers/_numba_adapter.py:138-160, kern/pic.py:164-205). */
int fb_pic_faces_to_particle(fb_ctx* ctx, int precision, int64_t n, const void* u, const void* v,
                             const void* positions, int nx, int ny, double x0, double y0, double dx, double dy,
                             int periodic_x, int periodic_y, void* out);
/* particle_to_grid (fb/solvers/_numba_adapter.py:35-71, kern/pic.py:22-65) on host arrays: positions and
 * velocities (n, 2), grid_out (ny, nx, 2) cell-centred; the stencil is wrapped on a periodic axis and clamped
 * to the edge cells otherwise; a cell of weight <= 1e-12 takes the freestream. Bit-identical to the
 * reference's particle-order accumulation. */
int fb_pic_particle_to_cells(fb_ctx* ctx, int precision, int64_t n, const void* positions, const void* velocities,
                             int nx, int ny, double x0, double y0, double dx, double dy, double freestream_x,
                             double freestream_y, int periodic_x, int periodic_y, void* grid_out);
/* grid_to_particle (fb/solvers/_numba_adapter.py:74-105, kern/pic.py:67-102): grid (ny, nx, 2), positions and
 * velocities_out (n, 2). */
int fb_pic_cells_to_particle(fb_ctx* ctx, int precision, int64_t n, const void* grid, const void* positions, int nx,
                             int ny, double x0, double y0, double dx, double dy, int periodic_x, int periodic_y,
                             void* velocities_out);

/* ----------------------------------------------------------------- Diagnostics ---- */
/* Run-level diagnostics computed on the device from the solver's fields (fb/core/metrics.py:24-104,
 * fb/solvers/stable_fluids.py:708-738, fb/solvers/lbm.py:1054-1146, fb/solvers/pic_flip.py:1115-1169), and
 * sample_velocity (fb/core/interpolation.py:9-79) of each solver. Per-cell terms are formed in the dtype of
 * the field the reference exports; counts and maxima equal the reference's on that field, means are FP64
 * sums in a fixed order. */
#define FB_DIAG_HISTOGRAM 1024

typedef struct fb_diag_params {
  double freestream_u;          /* scenario.freestream[0] */
  double wake_threshold;        /* 0.1 * |freestream_u| (fb/core/metrics.py:64-87) */
  int32_t wake_col0;            /* first column whose centre lies beyond pivot_x + chord (nx: none) */
  int32_t down_col0;            /* first column whose centre lies beyond pivot_x (nx: none) */
  double angular_velocity_rad;  /* lattice cut links: deg2rad of the control's angular velocity */
  double diagonal_link;         /* lattice cut links: hypot(dx, dy) */
  const double* centers_x;      /* lattice cut links: host, nx cell-centre x (linspace of the domain) */
  const double* centers_y;      /* lattice cut links: host, ny cell-centre y */
} fb_diag_params;

typedef struct fb_diag {
  double kinetic_energy;     /* 0.5 mean |u|^2 */
  double enstrophy;          /* 0.5 mean (dv/dx - du/dy)^2, np.gradient differences */
  double divergence_l2;      /* rms of du/dx + dv/dy */
  double wake_width;         /* rows with a deficit above the threshold in a wake column (fluid), times dy */
  double recirculation_area; /* fluid cells downstream of the pivot with u < 0, times dx dy */
  double density_mean;       /* lattice: mean density */
  double cut_link_adjacent_normal_speed; /* lattice: fb/solvers/lbm.py:1054-1101 */
  int64_t nonfinite;         /* non-finite cell velocity components */
  int64_t fluid_cells;       /* PIC/FLIP census over fluid cells (particle_cell_counts at the mask): */
  int64_t empty_cells;       /*   cells without particles */
  int64_t underfilled_cells; /*   cells with fewer than 4 */
  int64_t max_particles;     /*   largest count */
} fb_diag;

/* Stable Fluids cell velocity (face averages) and mask. */
int fb_mac_diagnostics(fb_mac* mac, const fb_diag_params* params, fb_diag* out);
/* Lattice physical velocity (lattice velocity times velocity_scale) and density, with cut-link speeds;
 * export_f64: the exported field is float64 (per-cell terms in FP64 on the lattice's FP32 values). */
int fb_lbm_diagnostics(fb_lbm* lbm, double velocity_scale, int export_f64, const fb_diag_params* params,
                       fb_diag* out);
/* PIC/FLIP grid velocity and particle census: histogram (FB_DIAG_HISTOGRAM int64, the last entry counting
 * every fluid cell at or above it) and, when counts is non-NULL, the (ny, nx) particle counts. */
int fb_pic_diagnostics(fb_pic* pic, const fb_diag_params* params, fb_diag* out, int64_t* histogram,
                       int32_t* counts);
/* Bilinear sample of the cell velocity at n points (n, 2) float64 into out (n, 2) of the exported dtype. */
int fb_mac_sample_velocity(fb_mac* mac, int64_t n, const double* points, void* out);
int fb_lbm_sample_velocity(fb_lbm* lbm, double velocity_scale, int export_f64, int64_t n, const double* points,
                           void* out);
int fb_pic_sample_velocity(fb_pic* pic, int64_t n, const double* points, void* out);

/* ------------------------------------------------------------------- Serving ------ */
/* No counterpart in the reference, which advances one scenario at a time in-process. These
 * entry points run many scenarios of uneven size concurrently on one GPU. */

/* Autotuner table: launch-shape choices that leave every result bit unchanged, per (parameter,
 * nx, ny, precision, grids); value -1 clears an entry, and unset entries use built-in defaults.
 * FB_TUNE_DST_ROWS: rows per block of the DST-I kernel (grids = batch size, 0 = any);
 * FB_TUNE_LBM_LIST_BLOCKS: blocks striding over a lattice's flagged cells (grids = 1);
 * FB_TUNE_LBM_FUSE: 1 runs eligible lattices (non-periodic, foil clear of the edges) two steps
 * per launch; otherwise (the default) one step per launch (grids = 1);
 * FB_TUNE_LBM_ROWS: interior rows per thread of the lattice step, 1 or 2; unset, 2 from 2 Mi
 * cells (grids = 1);
 * FB_TUNE_DST_PPT: points per thread of the DST-I kernel's FFT passes, 16 or 8; unset, 8 for
 * FFT lengths 512 to 4096 (grids as for FB_TUNE_DST_ROWS);
 * FB_TUNE_LBM_NT: 1 streams the interior populations of the lattice step with nontemporal loads and
 * stores, 0 not; unset, on once the two population copies exceed 192 MiB (grids = 1). */
enum fb_tune_param {
  FB_TUNE_DST_ROWS = 0,
  FB_TUNE_LBM_LIST_BLOCKS = 1,
  FB_TUNE_LBM_FUSE = 2,
  FB_TUNE_LBM_ROWS = 3,
  FB_TUNE_DST_PPT = 4,
  FB_TUNE_LBM_NT = 5
};
int fb_tune_set(int param, int nx, int ny, int precision, int grids, int value);
int fb_tune_get(int param, int nx, int ny, int precision, int grids, int* value);
/* Average milliseconds of one DST-I preconditioner application (FP32 or FP64) to `grids` stacked
 * nx x ny grids under the current tuning table, for the autotuner. */
int fb_poisson_time_precond(int device, int nx, int ny, double dx, double dy, int precision, int grids, int reps,
                            double* ms);

/* Device memory in bytes (hipMemGetInfo) for memory-based admission. */
int fb_mem_info(int device, int64_t* free_bytes, int64_t* total_bytes);

/* `steps` lattice steps of each of n lattices (one precision) in batched launches: one
 * launch per step covers every lattice, and chunks of `chunk` steps (even) are replayed
 * from one HIP graph. Same arithmetic per cell as fb_lbm_advance; stats[k] (may
 * be NULL) receives lattice k's postcondition reductions. Synchronous. */
int fb_lbm_run_batch(fb_lbm* const* lattices, const fb_lbm_params* params, int n, int steps, int chunk,
                     fb_lbm_stats* stats);

typedef struct fb_server fb_server;

typedef struct fb_server_options {
  int sf_workers;  /* Stable Fluids requests in flight at once, one stream each (>= 1) */
  int lbm_chunk;   /* lattice steps per graph launch (even, >= 2) */
  int use_graphs;  /* 1: replay HIP graphs (lattice batches, CG iteration chunks) */
  int sf_batch;    /* > 1: pressure solves of up to this many same-shape requests share launches */
  double sf_gather_us; /* how long a batched solve waits for more same-shape requests to join */
} fb_server_options;

enum fb_serve_kind { FB_SERVE_SF = 0, FB_SERVE_LBM = 1 };

typedef struct fb_server_result {
  int64_t id;
  int kind;               /* fb_serve_kind */
  int status;             /* FB_OK or the FB_ERR_* of the failing call */
  int sf_status;          /* fb_sf_status of the first substep that did not finish cleanly */
  int steps_done;
  double submitted_ms;    /* server clock, ms since fb_server_create */
  double started_ms;
  double finished_ms;
  int64_t pressure_iterations;  /* Stable Fluids: CG iterations over all substeps */
  fb_lbm_stats lbm_stats;       /* lattice: postcondition reductions of the final state */
} fb_server_result;

/* A request loop on `device`: Stable Fluids requests run on a pool of worker streams, lattice
 * requests join a continuous batch (one launch per step for all of them, graph-replayed in
 * chunks; lattices join and leave at chunk boundaries). Handles submitted to it belong to the
 * server until their result has been polled: the caller must not use them meanwhile. */
int fb_server_create(int device, const fb_server_options* options, fb_server** out);
/* Waits for the requests in flight, then stops the workers. */
int fb_server_destroy(fb_server* srv);
/* `steps` fixed substeps of params->dt at `pose` (each the body of fb_sf_substep). */
int fb_server_submit_sf(fb_server* srv, int64_t id, fb_mac* mac, const fb_sf_params* params, const fb_pose* pose,
                        int steps);
/* `steps` lattice steps with fixed parameters and pose. */
int fb_server_submit_lbm(fb_server* srv, int64_t id, fb_lbm* lbm, const fb_lbm_params* params, int steps);
/* Up to `max` finished requests, oldest first; *count receives how many were written. */
int fb_server_poll(fb_server* srv, fb_server_result* out, int max, int* count);
/* Requests queued (not yet started) and running. */
int fb_server_load(fb_server* srv, int* queued, int* running);
int fb_server_clock(fb_server* srv, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* FOILBENCH_AMD_H */
