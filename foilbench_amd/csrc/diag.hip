// Device diagnostics: run-level field metrics, lattice cut-link speed, particle census and velocity
// sampling (declared in diag.h), computed on the device from the fields where they live; only the few
// reduced numbers (or the sampled values) cross to the host.
//
// Every per-cell term is formed in the dtype of the field the reference's NumPy code sees (the FP32
// gradient quotient, the FP32 speed square, the FP32 comparison with the wake threshold), so counts and
// maxima agree bit for bit with the host metrics on the exported field; the means are sums in FP64 over
// rows in a fixed order (the reference sums in the field's dtype, pairwise).

#include <algorithm>
#include <cmath>
#include <vector>

#include "diag.h"

namespace fb {

#pragma clang fp contract(off)

namespace {

constexpr int kDiagThreads = 256;
constexpr int kRowSlots = 8;  // ke, enstrophy, div^2, rho, recirculating, wake row, cut-link max, non-finite

template <class S, class A>
__device__ __forceinline__ A cell_comp(const DiagField& f, const void* base, int comp, int x, int y) {
  const S* p = (const S*)base;
  if (f.kind == kDiagFaces) {
    // fb/solvers/stable_fluids.py cell_velocity: 0.5 * (u[:, :-1] + u[:, 1:]), 0.5 * (v[:-1] + v[1:])
    if (comp == 0) {
      const size_t r = (size_t)y * (f.nx + 1) + x;
      return (A)((A)((A)p[r] + (A)p[r + 1]) * (A)0.5);
    }
    const size_t r = (size_t)y * f.nx + x;
    return (A)((A)((A)p[r] + (A)p[r + f.nx]) * (A)0.5);
  }
  return (A)p[((size_t)y * f.nx + x) * f.stride];
}

template <class S, class A>
__device__ __forceinline__ A cu(const DiagField& f, int x, int y) {
  return cell_comp<S, A>(f, f.u, 0, x, y);
}
template <class S, class A>
__device__ __forceinline__ A cv(const DiagField& f, int x, int y) {
  return cell_comp<S, A>(f, f.v, 1, x, y);
}

// np.gradient with scalar spacing and edge_order 1 (first differences at the ends), in the dtype.
template <class S, class A, int COMP>
__device__ __forceinline__ A grad_x(const DiagField& f, int x, int y, A h, A h2) {
  auto val = [&](int xx) { return COMP ? cv<S, A>(f, xx, y) : cu<S, A>(f, xx, y); };
  if (x == 0) return (A)(val(1) - val(0)) / h;
  if (x == f.nx - 1) return (A)(val(f.nx - 1) - val(f.nx - 2)) / h;
  return (A)(val(x + 1) - val(x - 1)) / h2;
}
template <class S, class A, int COMP>
__device__ __forceinline__ A grad_y(const DiagField& f, int x, int y, A h, A h2) {
  auto val = [&](int yy) { return COMP ? cv<S, A>(f, x, yy) : cu<S, A>(f, x, yy); };
  if (y == 0) return (A)(val(1) - val(0)) / h;
  if (y == f.ny - 1) return (A)(val(f.ny - 1) - val(f.ny - 2)) / h;
  return (A)(val(y + 1) - val(y - 1)) / h2;
}

struct RowArgs {
  fb_diag_params p;
  const double* cx;  // cell centres (device), for cut links
  const double* cy;
};

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
  return v;
}

// One block per row: the row's sums, counts and maxima into part[y * kRowSlots + q].
template <class S, class A>
__global__ void __launch_bounds__(kDiagThreads) diag_row_kernel(DiagField f, RowArgs a, double* __restrict__ part) {
  const int y = blockIdx.x;
  const A hx = (A)f.dx, hy = (A)f.dy, h2x = (A)(2.0 * f.dx), h2y = (A)(2.0 * f.dy);
  const A fs = (A)a.p.freestream_u, thr = (A)a.p.wake_threshold;
  double ke = 0.0, ens = 0.0, d2 = 0.0, rho = 0.0, recirc = 0.0, cut = 0.0, bad = 0.0;
  int wake = 0;
  const int dxs[8] = {1, 0, -1, 0, 1, -1, -1, 1};
  const int dys[8] = {0, 1, 0, -1, 1, 1, -1, -1};
  for (int x = threadIdx.x; x < f.nx; x += kDiagThreads) {
    const size_t i = (size_t)y * f.nx + x;
    const A u = cu<S, A>(f, x, y), v = cv<S, A>(f, x, y);
    const A sp = (A)(u * u) + (A)(v * v);
    const A w = (A)(grad_x<S, A, 1>(f, x, y, hx, h2x) - grad_y<S, A, 0>(f, x, y, hy, h2y));
    const A dv = (A)(grad_x<S, A, 0>(f, x, y, hx, h2x) + grad_y<S, A, 1>(f, x, y, hy, h2y));
    ke += (double)sp;
    ens += (double)(A)(w * w);
    d2 += (double)(A)(dv * dv);
    if (!isfinite((double)u) || !isfinite((double)v)) bad += 1.0;
    const bool fluid = !f.solid || !f.solid[i];
    if (fluid && x >= a.p.down_col0 && u < (A)0) recirc += 1.0;
    if (fluid && x >= a.p.wake_col0 && (A)(fs - u) > thr) wake = 1;
    if (f.rho) rho += (double)(A)((const S*)f.rho)[i];
    if (f.sdf && fluid) {
      // fb/solvers/lbm.py:1054-1101: links from this fluid cell into a solid cell
      for (int d = 0; d < 8; ++d) {
        int xd = x + dxs[d], yd = y + dys[d];
        if (f.px) xd = wrapi(xd, f.nx);
        else if (xd < 0 || xd >= f.nx) continue;
        if (f.py) yd = wrapi(yd, f.ny);
        else if (yd < 0 || yd >= f.ny) continue;
        const size_t j = (size_t)yd * f.nx + xd;
        if (!f.solid[j]) continue;
        const double s0 = f.sdf[i];
        double fr = s0 / fmax(s0 - f.sdf[j], 1.0e-12);
        fr = fmin(fmax(fr, 0.05), 1.0);
        const double lx = (double)dxs[d] * f.dx, ly = (double)dys[d] * f.dy;
        const double px = a.cx[x] + fr * lx, py = a.cy[y] + fr * ly;
        const double rx = px - f.pivot_x, ry = py - f.pivot_y;
        const double om = a.p.angular_velocity_rad;
        const double relx = (double)u - (-om * ry), rely = (double)v - om * rx;
        const double len = (dxs[d] != 0 && dys[d] != 0) ? a.p.diagonal_link : (dxs[d] != 0 ? f.dx : f.dy);
        const double proj = fabs(relx * lx + rely * ly) / len;
        cut = nan_max(cut, proj);
      }
    }
  }
  double vals[kRowSlots] = {ke, ens, d2, rho, recirc, (double)wake, cut, bad};
  __shared__ double lds[kDiagThreads / 64][kRowSlots];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kRowSlots; ++q) {
    double r = (q == 5 || q == 6) ? wave_max_d(vals[q]) : wave_sum_d(vals[q]);
    if (lane == 0) lds[wv][q] = r;
  }
  __syncthreads();
  if (threadIdx.x < kRowSlots) {
    const int q = threadIdx.x;
    double r = lds[0][q];
    for (int k = 1; k < kDiagThreads / 64; ++k) r = (q == 5 || q == 6) ? nan_max(r, lds[k][q]) : r + lds[k][q];
    part[(size_t)y * kRowSlots + q] = r;
  }
}

// Rows folded in a fixed order: thread t takes rows t, t + 256, ..., then a fixed tree.
__global__ void __launch_bounds__(kDiagThreads) diag_fold_kernel(int ny, const double* __restrict__ part,
                                                                 double* __restrict__ out) {
  __shared__ double lds[kDiagThreads][kRowSlots];
  double acc[kRowSlots];
#pragma unroll
  for (int q = 0; q < kRowSlots; ++q) acc[q] = 0.0;
  for (int y = threadIdx.x; y < ny; y += kDiagThreads) {
#pragma unroll
    for (int q = 0; q < kRowSlots; ++q) {
      const double v = part[(size_t)y * kRowSlots + q];
      acc[q] = q == 6 ? nan_max(acc[q], v) : acc[q] + v;
    }
  }
#pragma unroll
  for (int q = 0; q < kRowSlots; ++q) lds[threadIdx.x][q] = acc[q];
  __syncthreads();
  for (int s = kDiagThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < kRowSlots; ++q) {
        const double o = lds[threadIdx.x + s][q];
        lds[threadIdx.x][q] = q == 6 ? nan_max(lds[threadIdx.x][q], o) : lds[threadIdx.x][q] + o;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < kRowSlots) out[threadIdx.x] = lds[0][threadIdx.x];
}

// Python's float modulo for a positive period (numpy's npy_divmod).
__device__ __forceinline__ double pymod(double a, double b) {
  double m = fmod(a, b);
  if (m != 0.0 && m < 0.0) m += b;
  return m;
}

// fb/core/interpolation.py:9-52 with offset (0.5, 0.5) on each velocity component.
template <class S, class A>
__global__ void __launch_bounds__(256) sample_kernel(DiagField f, int n, const double* __restrict__ pts,
                                                     A* __restrict__ out) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  double gx = (pts[2 * k] - f.x0) / f.dx - 0.5;
  double gy = (pts[2 * k + 1] - f.y0) / f.dy - 0.5;
  gx = f.px ? pymod(gx, (double)f.nx) : fmin(fmax(gx, 0.0), f.nx - 1.0);
  gy = f.py ? pymod(gy, (double)f.ny) : fmin(fmax(gy, 0.0), f.ny - 1.0);
  const double fx = floor(gx), fy = floor(gy);
  // a NaN coordinate gives no valid index: clamp so the loads stay inside the field
  int ix0 = (fx >= 0.0 && fx < (double)f.nx) ? (int)fx : (fx >= (double)f.nx ? f.nx - 1 : 0);
  int iy0 = (fy >= 0.0 && fy < (double)f.ny) ? (int)fy : (fy >= (double)f.ny ? f.ny - 1 : 0);
  const double tx = gx - (double)ix0, ty = gy - (double)iy0;
  const int ix1 = f.px ? (ix0 + 1) % f.nx : min(ix0 + 1, f.nx - 1);
  const int iy1 = f.py ? (iy0 + 1) % f.ny : min(iy0 + 1, f.ny - 1);
  for (int c = 0; c < 2; ++c) {
    const void* base = c ? f.v : f.u;
    const double a00 = (double)cell_comp<S, A>(f, base, c, ix0, iy0);
    const double a01 = (double)cell_comp<S, A>(f, base, c, ix1, iy0);
    const double a10 = (double)cell_comp<S, A>(f, base, c, ix0, iy1);
    const double a11 = (double)cell_comp<S, A>(f, base, c, ix1, iy1);
    const double r = (1.0 - tx) * (1.0 - ty) * a00 + tx * (1.0 - ty) * a01 + (1.0 - tx) * ty * a10 + tx * ty * a11;
    out[2 * k + c] = (A)r;
  }
}

// Particles per cell (fb/core/particle_population.py:19-31): floor((x - x0) / dx) in the dtype, clipped.
template <class T>
__global__ void __launch_bounds__(256) census_count_kernel(int n, const T* __restrict__ pos, T x0, T y0, T dx, T dy,
                                                           int nx, int ny, int* __restrict__ cnt) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const T x = pos[2 * k], y = pos[2 * k + 1];
  const double fx = (double)floor((T)((T)(x - x0) / dx));
  const double fy = (double)floor((T)((T)(y - y0) / dy));
  const int cx = fx >= 0.0 ? (fx <= (double)(nx - 1) ? (int)fx : nx - 1) : 0;
  const int cy = fy >= 0.0 ? (fy <= (double)(ny - 1) ? (int)fy : ny - 1) : 0;
  atomicAdd(&cnt[(size_t)cy * nx + cx], 1);
}

// Histogram of the fluid cells' counts, staged per block in LDS.
__global__ void __launch_bounds__(256) census_hist_kernel(int ncell, const int* __restrict__ cnt,
                                                          const uint8_t* __restrict__ solid,
                                                          unsigned long long* __restrict__ hist,
                                                          int* __restrict__ maxc) {
  __shared__ unsigned int h[FB_DIAG_HISTOGRAM];
  for (int b = threadIdx.x; b < FB_DIAG_HISTOGRAM; b += 256) h[b] = 0u;
  __syncthreads();
  int m = -1;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < ncell; i += gridDim.x * 256) {
    if (solid[i]) continue;
    const int c = cnt[i];
    atomicAdd(&h[min(c, FB_DIAG_HISTOGRAM - 1)], 1u);
    m = max(m, c);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0 && m >= 0) atomicMax(maxc, m);
  __syncthreads();
  for (int b = threadIdx.x; b < FB_DIAG_HISTOGRAM; b += 256)
    if (h[b]) atomicAdd(&hist[b], (unsigned long long)h[b]);
}

__global__ void zero_kernel(unsigned char* __restrict__ p, size_t bytes) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < bytes; i += (size_t)gridDim.x * 256) p[i] = 0;
}

void zero_async(hipStream_t s, void* p, size_t bytes) {
  const unsigned blocks = (unsigned)std::min<size_t>(1024, (bytes + 255) / 256 + 1);
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, s, (unsigned char*)p, bytes);
}

template <class S, class A>
int run_rows(Ctx* c, const DiagField& f, const RowArgs& a, double* part, double* out) {
  hipLaunchKernelGGL((diag_row_kernel<S, A>), dim3(f.ny), dim3(kDiagThreads), 0, c->stream, f, a, part);
  hipLaunchKernelGGL(diag_fold_kernel, dim3(1), dim3(kDiagThreads), 0, c->stream, f.ny, part, out);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

template <class S, class A>
int run_sample(Ctx* c, const DiagField& f, int n, const double* pts, void* out) {
  hipLaunchKernelGGL((sample_kernel<S, A>), dim3((n + 255) / 256), dim3(256), 0, c->stream, f, n, pts, (A*)out);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

int check_field(const DiagField& f) {
  FB_REQUIRE(f.nx >= 2 && f.ny >= 2, FB_ERR_ARG, "diagnostics need at least two cells per axis");
  FB_REQUIRE(f.u && f.v, FB_ERR_ARG, "NULL velocity field");
  FB_REQUIRE(f.storage_f64 || !f.arith_f64 || f.kind == kDiagCells, FB_ERR_ARG, "unsupported field form");
  FB_REQUIRE(!f.storage_f64 || f.arith_f64, FB_ERR_ARG, "an FP64 field is exported in FP64");
  return FB_OK;
}

}  // namespace

int field_diagnostics(Ctx* c, const DiagField& f, const fb_diag_params& p, fb_diag* out) {
  int rc = check_field(f);
  if (rc) return rc;
  FB_REQUIRE(!f.sdf || (f.solid && p.centers_x && p.centers_y), FB_ERR_ARG,
             "cut-link speeds need the mask and the cell centres");
  hipStream_t s = c->stream;
  const size_t part_bytes = (size_t)f.ny * kRowSlots * sizeof(double);
  double *part = nullptr, *dout = nullptr, *cx = nullptr, *cy = nullptr;
  hipError_t e = hipMallocAsync((void**)&part, part_bytes, s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&dout, kRowSlots * sizeof(double), s);
  if (e == hipSuccess && f.sdf) e = hipMallocAsync((void**)&cx, (size_t)f.nx * sizeof(double), s);
  if (e == hipSuccess && f.sdf) e = hipMallocAsync((void**)&cy, (size_t)f.ny * sizeof(double), s);
  if (e == hipSuccess && f.sdf)
    e = hipMemcpyAsync(cx, p.centers_x, (size_t)f.nx * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && f.sdf)
    e = hipMemcpyAsync(cy, p.centers_y, (size_t)f.ny * sizeof(double), hipMemcpyHostToDevice, s);
  double host[kRowSlots] = {};
  if (e == hipSuccess) {
    RowArgs a{p, cx, cy};
    if (f.storage_f64) rc = run_rows<double, double>(c, f, a, part, dout);
    else if (f.arith_f64) rc = run_rows<float, double>(c, f, a, part, dout);
    else rc = run_rows<float, float>(c, f, a, part, dout);
    if (!rc) e = hipMemcpyAsync(host, dout, sizeof(host), hipMemcpyDeviceToHost, s);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(s);
  }
  for (void* q : {(void*)part, (void*)dout, (void*)cx, (void*)cy})
    if (q) (void)hipFreeAsync(q, s);
  if (e != hipSuccess) {
    set_error("diagnostics failed: %s", hipGetErrorString(e));
    return FB_ERR_HIP;
  }
  if (rc) return rc;
  const double n = (double)f.nx * (double)f.ny;
  // fb/core/metrics.py:30-31, 46-48, 51-55, 64-87, 90-104
  out->kinetic_energy = 0.5 * (host[0] / n);
  out->enstrophy = 0.5 * (host[1] / n);
  out->divergence_l2 = std::sqrt(host[2] / n);
  out->density_mean = f.rho ? host[3] / n : 0.0;
  out->recirculation_area = host[4] * f.dx * f.dy;
  out->wake_width = host[5] * f.dy;
  out->cut_link_adjacent_normal_speed = host[6];
  out->nonfinite = (int64_t)host[7];
  FB_HIP(hipStreamSynchronize(s));
  return FB_OK;
}

int sample_cell_velocity(Ctx* c, const DiagField& f, int64_t n, const double* points_host, void* out_host) {
  int rc = check_field(f);
  if (rc) return rc;
  FB_REQUIRE(n >= 0 && n < (1ll << 28), FB_ERR_ARG, "point count out of range");
  if (n == 0) return FB_OK;
  FB_REQUIRE(points_host && out_host, FB_ERR_ARG, "NULL point arrays");
  hipStream_t s = c->stream;
  const size_t ob = (size_t)n * 2 * (f.arith_f64 ? sizeof(double) : sizeof(float));
  double* pts = nullptr;
  void* dout = nullptr;
  hipError_t e = hipMallocAsync((void**)&pts, (size_t)n * 2 * sizeof(double), s);
  if (e == hipSuccess) e = hipMallocAsync(&dout, ob, s);
  if (e == hipSuccess) e = hipMemcpyAsync(pts, points_host, (size_t)n * 2 * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    if (f.storage_f64) rc = run_sample<double, double>(c, f, (int)n, pts, dout);
    else if (f.arith_f64) rc = run_sample<float, double>(c, f, (int)n, pts, dout);
    else rc = run_sample<float, float>(c, f, (int)n, pts, dout);
    if (!rc) e = hipMemcpyAsync(out_host, dout, ob, hipMemcpyDeviceToHost, s);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(s);
  }
  for (void* q : {(void*)pts, dout})
    if (q) (void)hipFreeAsync(q, s);
  if (e != hipSuccess) {
    set_error("velocity sampling failed: %s", hipGetErrorString(e));
    return FB_ERR_HIP;
  }
  FB_HIP(hipStreamSynchronize(s));
  return rc;
}

int particle_census(Ctx* c, const DiagField& f, int64_t n, const void* positions, fb_diag* out, int64_t* hist_host,
                    int32_t* counts_host) {
  FB_REQUIRE(f.solid && hist_host, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(n == 0 || positions, FB_ERR_ARG, "NULL positions");
  hipStream_t s = c->stream;
  const int ncell = f.nx * f.ny;
  int* cnt = nullptr;
  unsigned long long* hist = nullptr;
  int* maxc = nullptr;
  hipError_t e = hipMallocAsync((void**)&cnt, (size_t)ncell * sizeof(int), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&hist, FB_DIAG_HISTOGRAM * sizeof(unsigned long long) + 16, s);
  if (e == hipSuccess) {
    maxc = reinterpret_cast<int*>(hist + FB_DIAG_HISTOGRAM);
    zero_async(s, cnt, (size_t)ncell * sizeof(int));
    zero_async(s, hist, FB_DIAG_HISTOGRAM * sizeof(unsigned long long) + 16);
    if (n > 0) {
      const unsigned b = (unsigned)((n + 255) / 256);
      if (f.storage_f64)
        hipLaunchKernelGGL(census_count_kernel<double>, dim3(b), dim3(256), 0, s, (int)n, (const double*)positions,
                           f.x0, f.y0, f.dx, f.dy, f.nx, f.ny, cnt);
      else
        hipLaunchKernelGGL(census_count_kernel<float>, dim3(b), dim3(256), 0, s, (int)n, (const float*)positions,
                           (float)f.x0, (float)f.y0, (float)f.dx, (float)f.dy, f.nx, f.ny, cnt);
    }
    const unsigned hb = (unsigned)std::min(1024, (ncell + 255) / 256);
    hipLaunchKernelGGL(census_hist_kernel, dim3(hb), dim3(256), 0, s, ncell, cnt, f.solid, hist, maxc);
    e = hipGetLastError();
  }
  std::vector<unsigned long long> h(FB_DIAG_HISTOGRAM + 2, 0ull);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), hist, FB_DIAG_HISTOGRAM * sizeof(unsigned long long) + 16, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && counts_host)
    e = hipMemcpyAsync(counts_host, cnt, (size_t)ncell * sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (void* q : {(void*)cnt, (void*)hist})
    if (q) (void)hipFreeAsync(q, s);
  if (e != hipSuccess) {
    set_error("particle census failed: %s", hipGetErrorString(e));
    return FB_ERR_HIP;
  }
  int64_t fluid = 0, under = 0;
  for (int b = 0; b < FB_DIAG_HISTOGRAM; ++b) {
    hist_host[b] = (int64_t)h[b];
    fluid += (int64_t)h[b];
    if (b < 4) under += (int64_t)h[b];
  }
  int mx = 0;
  memcpy(&mx, &h[FB_DIAG_HISTOGRAM], sizeof(int));
  out->fluid_cells = fluid;
  out->empty_cells = (int64_t)h[0];
  out->underfilled_cells = under;
  out->max_particles = fluid ? mx : 0;
  FB_HIP(hipStreamSynchronize(s));
  return FB_OK;
}

}  // namespace fb
