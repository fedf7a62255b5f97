/*
 * D2Q9 TRT lattice steps in plain C: a restatement of oracle/lbm.py kernel_steps, itself restating
 * kern/lbm.py:12-168 (trt_collision_kernel, moving_wall_stream_kernel, sponge_kernel) and the boundary
 * treatment of fb/solvers/lbm.py:596-650 (Zou-He rows and inlet, convective outlet, corner equilibria).
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): built by oracle/build.py into oracle/liblbm_oracle.so and used by
 * tests/ as the checker of the GPU lattice at full configuration sizes and horizons, and by bench.py as the
 * CPU baseline (C loops in place of the reference's numba kernels). Never linked into the product.
 *
 * Operation order follows the NumPy restatement exactly, so results are bit-identical to it:
 * collision, streaming/Bouzidi reflection and sponge in FP64 arithmetic, rounded on store to the storage
 * type; Zou-He, outlet and corner boundaries in storage-type arithmetic. Compile with -ffp-contract=off.
 * Rows are independent within each stage, so the OpenMP loops over rows give the same bits for any thread
 * count.
 *
 * Layout: populations f (ny, nx, 9) AoS in the storage type, outlet history (ny, 9), solid (ny, nx) uint8,
 * sdf (ny, nx) double, sponge strength (ny, nx) double.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const double W[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                            1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

typedef struct lbm_oracle_params {
  int nx, ny;
  int periodic_x, periodic_y, channel;
  double omega_plus, omega_minus;
  double u_lattice;          /* outlet speed */
  double target_ux, target_uy;
  double feq_b[9];           /* boundary equilibrium, already in the storage type's values */
  double x0, y0, dx, dy, pivot_x, pivot_y;
  double omega_rad, scale;   /* wall angular velocity (rad/s) and the lattice/physical speed ratio */
} lbm_oracle_params;

static inline int wrap(int i, int n) { return ((i % n) + n) % n; }

/* One stage body per storage type, generated from a macro so the float and double paths share text. */
#define LBM_ORACLE_STEP(NAME, T)                                                                                      \
  static void NAME(const lbm_oracle_params* p, T* f, T* outlet, const uint8_t* solid, const double* sdf,             \
                   const double* sponge, T* post, T* rho_s, T* out, T* newout) {                                   \
    const int nx = p->nx, ny = p->ny;                                                                              \
    const double even = 0.5 * (p->omega_plus + p->omega_minus);                                                    \
    const double odd = 0.5 * (p->omega_plus - p->omega_minus);                                                     \
    /* trt_collision (oracle/lbm.py trt_collision): FP64, rounded to T */                                          \
    _Pragma("omp parallel for schedule(static)") for (int y = 0; y < ny; ++y) {                                    \
      for (int x = 0; x < nx; ++x) {                                                                               \
        const size_t i = (size_t)y * nx + x;                                                                       \
        double g[9], feq[9];                                                                                       \
        for (int d = 0; d < 9; ++d) g[d] = (double)f[i * 9 + d];                                                   \
        double rho = g[0];                                                                                         \
        for (int d = 1; d < 9; ++d) rho = rho + g[d];                                                              \
        const double safe = rho < 1.0e-12 ? 1.0e-12 : rho; /* np.maximum: NaN stays NaN */                       \
        double ux = ((((g[1] - g[3]) + g[5]) - g[6]) - g[7]) + g[8];                                               \
        ux = ux / safe;                                                                                            \
        double uy = ((((g[2] - g[4]) + g[5]) + g[6]) - g[7]) - g[8];                                               \
        uy = uy / safe;                                                                                            \
        const double s2 = ux * ux + uy * uy;                                                                       \
        for (int d = 0; d < 9; ++d) {                                                                              \
          const double pr = ((double)CX[d] * ux) + ((double)CY[d] * uy);                                           \
          feq[d] = (rho * W[d]) * (((1.0 + 3.0 * pr) + (4.5 * pr) * pr) - 1.5 * s2);                                \
        }                                                                                                          \
        for (int d = 0; d < 9; ++d) {                                                                              \
          const int o = OPP[d];                                                                                    \
          post[i * 9 + d] = (T)((g[d] - even * (g[d] - feq[d])) - odd * (g[o] - feq[o]));                          \
        }                                                                                                          \
        rho_s[i] = (T)rho;                                                                                         \
      }                                                                                                            \
    }                                                                                                              \
    /* moving_wall_stream: push streaming, Bouzidi on cut links (out starts at zero, +=) */                       \
    _Pragma("omp parallel for schedule(static)") for (int y = 0; y < ny; ++y) {                                    \
      for (int x = 0; x < nx; ++x) {                                                                               \
        const size_t i = (size_t)y * nx + x;                                                                       \
        out[i * 9 + 0] = post[i * 9 + 0];                                                                          \
        for (int d = 1; d < 9; ++d) {                                                                              \
          /* the value arriving at (x, y) along d comes from (x - c, y - c) when that cell is fluid and its   */ \
          /* destination (x, y) is fluid */                                                                        \
          const int sx = wrap(x - CX[d], nx), sy = wrap(y - CY[d], ny);                                            \
          const size_t s = (size_t)sy * nx + sx;                                                                   \
          const T moved = (!solid[s] && !solid[i]) ? post[s * 9 + d] : (T)0;                                       \
          out[i * 9 + d] = (T)0 + moved;                                                                           \
        }                                                                                                          \
      }                                                                                                            \
    }                                                                                                              \
    _Pragma("omp parallel for schedule(static)") for (int y = 0; y < ny; ++y) {                                    \
      for (int x = 0; x < nx; ++x) {                                                                               \
        const size_t i = (size_t)y * nx + x;                                                                       \
        if (solid[i]) continue;                                                                                    \
        for (int d = 1; d < 9; ++d) {                                                                              \
          const int tx = wrap(x + CX[d], nx), ty = wrap(y + CY[d], ny);                                            \
          const size_t t = (size_t)ty * nx + tx;                                                                   \
          if (!solid[t]) continue;                                                                                 \
          const double sdf_n = sdf[t];                                                                             \
          double denom = sdf[i] - sdf_n;                                                                           \
          if (denom < 1.0e-12) denom = 1.0e-12;                                                                    \
          double q = sdf[i] / denom;                                                                               \
          if (q < 0.05) q = 0.05;                                                                                  \
          if (q > 1.0) q = 1.0;                                                                                    \
          const int o = OPP[d];                                                                                    \
          const double src = (double)post[i * 9 + d];                                                              \
          double refl;                                                                                             \
          if (q < 0.5) {                                                                                           \
            const int ux_ = wrap(x - CX[d], nx), uy_ = wrap(y - CY[d], ny);                                        \
            const double up = (double)post[((size_t)uy_ * nx + ux_) * 9 + d];                                      \
            refl = (2.0 * q) * src + (1.0 - 2.0 * q) * up;                                                         \
          } else {                                                                                                 \
            refl = src / (2.0 * q) + ((2.0 * q - 1.0) * (double)post[i * 9 + o]) / (2.0 * q);                      \
          }                                                                                                        \
          const double wx = (p->x0 + ((double)x + 0.5) * p->dx) + (q * (double)CX[d]) * p->dx;                     \
          const double wy = (p->y0 + ((double)y + 0.5) * p->dy) + (q * (double)CY[d]) * p->dy;                     \
          const double wu = (-p->omega_rad) * (wy - p->pivot_y);                                                   \
          const double wv = p->omega_rad * (wx - p->pivot_x);                                                      \
          const double wp = ((double)CX[d] * wu) + ((double)CY[d] * wv);                                           \
          refl = refl - (((6.0 * W[d]) * (double)rho_s[i]) * wp) * p->scale;                                       \
          /* the reflected value replaces out[x, opp(d)]; a later direction's push into it adds a zero */          \
          out[i * 9 + o] = (T)refl;                                                                                \
        }                                                                                                          \
      }                                                                                                            \
    }                                                                                                              \
    /* a push along o into a cut-link cell comes from its solid neighbour and is zero: out += 0 */                 \
    _Pragma("omp parallel for schedule(static)") for (int y = 0; y < ny; ++y) {                                    \
      for (int x = 0; x < nx; ++x) {                                                                               \
        const size_t i = (size_t)y * nx + x;                                                                       \
        if (solid[i]) continue;                                                                                    \
        for (int d = 1; d < 9; ++d) {                                                                              \
          const int tx = wrap(x + CX[d], nx), ty = wrap(y + CY[d], ny);                                            \
          if (!solid[(size_t)ty * nx + tx]) continue;                                                              \
          const int o = OPP[d];                                                                                    \
          if (o > d) out[i * 9 + o] = out[i * 9 + o] + (T)0;                                                       \
        }                                                                                                          \
      }                                                                                                            \
    }                                                                                                              \
    /* boundaries in storage arithmetic (oracle/lbm.py bottom/top/left_velocity) */                               \
    const T tux = (T)p->target_ux, tuy = (T)p->target_uy;                                                          \
    if (!p->periodic_y) {                                                                                          \
      if (p->channel) {                                                                                            \
        for (int x = 0; x < nx; ++x) {                                                                             \
          T a[9], b[9];                                                                                            \
          for (int d = 0; d < 9; ++d) {                                                                            \
            a[d] = out[((size_t)1 * nx + x) * 9 + OPP[d]];                                                         \
            b[d] = out[((size_t)(ny - 2) * nx + x) * 9 + OPP[d]];                                                  \
          }                                                                                                        \
          for (int d = 0; d < 9; ++d) {                                                                            \
            out[(size_t)x * 9 + d] = a[d];                                                                         \
            out[((size_t)(ny - 1) * nx + x) * 9 + d] = b[d];                                                       \
          }                                                                                                        \
        }                                                                                                          \
      } else {                                                                                                     \
        const T den_b = (T)(1.0 - p->target_uy), den_t = (T)(1.0 + p->target_uy);                                  \
        for (int x = 0; x < nx; ++x) {                                                                             \
          T* b = out + (size_t)x * 9;                                                                              \
          const T rho = (((b[0] + b[1]) + b[3]) + (T)2.0 * ((b[4] + b[7]) + b[8])) / den_b;                        \
          b[2] = b[4] + ((T)(2.0 / 3.0) * rho) * tuy;                                                              \
          b[5] = ((b[7] + (T)0.5 * (b[3] - b[1])) + ((T)(1.0 / 6.0) * rho) * tuy) + ((T)0.5 * rho) * tux;          \
          b[6] = ((b[8] + (T)0.5 * (b[1] - b[3])) + ((T)(1.0 / 6.0) * rho) * tuy) - ((T)0.5 * rho) * tux;          \
        }                                                                                                          \
        for (int x = 0; x < nx; ++x) {                                                                             \
          T* b = out + ((size_t)(ny - 1) * nx + x) * 9;                                                            \
          const T rho = (((b[0] + b[1]) + b[3]) + (T)2.0 * ((b[2] + b[5]) + b[6])) / den_t;                        \
          b[4] = b[2] - ((T)(2.0 / 3.0) * rho) * tuy;                                                              \
          b[7] = ((b[5] + (T)0.5 * (b[1] - b[3])) - ((T)(1.0 / 6.0) * rho) * tuy) - ((T)0.5 * rho) * tux;          \
          b[8] = ((b[6] + (T)0.5 * (b[3] - b[1])) - ((T)(1.0 / 6.0) * rho) * tuy) + ((T)0.5 * rho) * tux;          \
        }                                                                                                          \
      }                                                                                                            \
    }                                                                                                              \
    if (!p->periodic_x) {                                                                                          \
      const T den_l = (T)(1.0 - p->target_ux);                                                                     \
      for (int y = 0; y < ny; ++y) {                                                                               \
        T* b = out + ((size_t)y * nx) * 9;                                                                         \
        const T rho = (((b[0] + b[2]) + b[4]) + (T)2.0 * ((b[3] + b[6]) + b[7])) / den_l;                          \
        b[1] = b[3] + ((T)(2.0 / 3.0) * rho) * tux;                                                                \
        b[5] = ((b[7] + (T)0.5 * (b[4] - b[2])) + ((T)(1.0 / 6.0) * rho) * tux) + ((T)0.5 * rho) * tuy;            \
        b[8] = ((b[6] + (T)0.5 * (b[2] - b[4])) + ((T)(1.0 / 6.0) * rho) * tux) - ((T)0.5 * rho) * tuy;            \
      }                                                                                                            \
      const T u = (T)p->u_lattice;                                                                                 \
      for (int y = 0; y < ny; ++y) {                                                                               \
        for (int d = 0; d < 9; ++d) {                                                                              \
          const T prev = outlet[(size_t)y * 9 + d];                                                                \
          const T v = prev + u * (out[((size_t)y * nx + nx - 2) * 9 + d] - prev);                                 \
          out[((size_t)y * nx + nx - 1) * 9 + d] = v;                                                              \
          newout[(size_t)y * 9 + d] = v;                                                                           \
        }                                                                                                          \
      }                                                                                                            \
      memcpy(outlet, newout, (size_t)ny * 9 * sizeof(T));                                                          \
    }                                                                                                              \
    if (!p->periodic_x && !p->periodic_y) {                                                                        \
      const size_t corners[4] = {0, (size_t)(ny - 1) * nx, (size_t)nx - 1, (size_t)(ny - 1) * nx + nx - 1};       \
      for (int k = 0; k < 4; ++k)                                                                                  \
        for (int d = 0; d < 9; ++d) out[corners[k] * 9 + d] = (T)p->feq_b[d];                                      \
    }                                                                                                              \
    /* sponge (oracle/lbm.py sponge): FP64 mix where the strength is non-zero */                                  \
    _Pragma("omp parallel for schedule(static)") for (int y = 0; y < ny; ++y) {                                    \
      for (int x = 0; x < nx; ++x) {                                                                               \
        const size_t i = (size_t)y * nx + x;                                                                       \
        const double s = sponge[i];                                                                                \
        for (int d = 0; d < 9; ++d) {                                                                              \
          const T v = out[i * 9 + d];                                                                              \
          f[i * 9 + d] = s != 0.0 ? (T)((1.0 - s) * (double)v + s * (double)(T)p->feq_b[d]) : v;                  \
        }                                                                                                          \
      }                                                                                                            \
    }                                                                                                              \
  }

LBM_ORACLE_STEP(step_f64, double)
LBM_ORACLE_STEP(step_f32, float)

/* Runs `steps` lattice steps in place on f and outlet. storage: 8 (double) or 4 (float). threads <= 0 keeps
 * the OpenMP default. Returns 0, or 1 on a bad argument, 2 when scratch memory is unavailable. */
int lbm_oracle_steps(const lbm_oracle_params* p, int storage, void* f, void* outlet, const uint8_t* solid,
                  const double* sdf, const double* sponge, int steps, int threads) {
  if (!p || !f || !outlet || !solid || !sdf || !sponge || p->nx < 3 || p->ny < 3 || steps < 0) return 1;
  if (storage != 4 && storage != 8) return 1;
#ifdef _OPENMP
  extern void omp_set_num_threads(int);
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  const size_t n = (size_t)p->nx * p->ny;
  void* post = malloc(n * 9 * storage);
  void* rho = malloc(n * storage);
  void* out = malloc(n * 9 * storage);
  void* newout = malloc((size_t)p->ny * 9 * storage);
  if (!post || !rho || !out || !newout) {
    free(post);
    free(rho);
    free(out);
    free(newout);
    return 2;
  }
  for (int k = 0; k < steps; ++k) {
    if (storage == 8)
      step_f64(p, (double*)f, (double*)outlet, solid, sdf, sponge, (double*)post, (double*)rho, (double*)out,
               (double*)newout);
    else
      step_f32(p, (float*)f, (float*)outlet, solid, sdf, sponge, (float*)post, (float*)rho, (float*)out,
               (float*)newout);
  }
  free(post);
  free(rho);
  free(out);
  free(newout);
  return 0;
}
