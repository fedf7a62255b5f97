// Geometry kernels: NACA signed distance of cell centres and of arbitrary points.
#include "geometry.h"

namespace fb {

__global__ void __launch_bounds__(256) cell_sdf_kernel(CellGrid cg, NacaDev g, double* __restrict__ sdf,
                                                       uint8_t* __restrict__ solid) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= cg.nx || y >= cg.ny) return;
  double px = linspace_at(x, cg.nx, cg.sx, cg.stx, cg.ex);
  double py = linspace_at(y, cg.ny, cg.sy, cg.sty, cg.ey);
  double d = naca_sdf(g, px, py);
  size_t i = (size_t)y * cg.nx + x;
  sdf[i] = d;
  solid[i] = (d <= 0.0) ? 1 : 0;
}

int launch_cell_sdf(hipStream_t s, const CellGrid& cg, const NacaDev& g, double* sdf, uint8_t* solid) {
  dim3 block(256);
  dim3 grid((cg.nx + 255) / 256, cg.ny);
  hipLaunchKernelGGL(cell_sdf_kernel, grid, block, 0, s, cg, g, sdf, solid);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

#pragma clang fp contract(off)
// normals (fb/core/geometry.py:108-126): centred differences of the distance.
__global__ void __launch_bounds__(256) point_sdf_kernel(NacaDev g, int64_t n, const double* __restrict__ pts,
                                                        double* __restrict__ sdf, double* __restrict__ nrm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double px = pts[2 * i], py = pts[2 * i + 1];
  sdf[i] = naca_sdf(g, px, py);
  if (nrm) {
    double eps = fmax(g.chord * 1.0e-4, 1.0e-6);
    double gx = naca_sdf(g, px + eps, py + 0.0) - naca_sdf(g, px - eps, py - 0.0);
    double gy = naca_sdf(g, px + 0.0, py + eps) - naca_sdf(g, px - 0.0, py - eps);
    double len = sqrt(gx * gx + gy * gy);
    if (len < eps) {
      gx = -g.sin_a;
      gy = g.cos_a;
      len = 1.0;
    }
    double den = fmax(len, eps);
    nrm[2 * i] = gx / den;
    nrm[2 * i + 1] = gy / den;
  }
}
#pragma clang fp contract(fast)

}  // namespace fb

extern "C" {

int fb_naca_cell_sdf(fb_ctx* ctx, const fb_domain* dom, const fb_foil* foil, const fb_pose* pose,
                     double* sdf_host, uint8_t* solid_host) {
  FB_REQUIRE(ctx && dom && foil && pose, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(dom->nx >= 2 && dom->ny >= 2, FB_ERR_ARG, "grid too small");
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  size_t n = (size_t)dom->nx * dom->ny;
  double* d_sdf = nullptr;
  uint8_t* d_solid = nullptr;
  FB_HIP(hipMallocAsync((void**)&d_sdf, n * sizeof(double), c->stream));
  FB_HIP(hipMallocAsync((void**)&d_solid, n, c->stream));
  int rc = fb::launch_cell_sdf(c->stream, fb::make_cell_grid(*dom), fb::make_naca(*foil, *pose), d_sdf, d_solid);
  if (rc == FB_OK) {
    if (sdf_host) FB_HIP(hipMemcpyAsync(sdf_host, d_sdf, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (solid_host) FB_HIP(hipMemcpyAsync(solid_host, d_solid, n, hipMemcpyDeviceToHost, c->stream));
  }
  FB_HIP(hipFreeAsync(d_sdf, c->stream));
  FB_HIP(hipFreeAsync(d_solid, c->stream));
  FB_HIP(hipStreamSynchronize(c->stream));
  return rc;
}

int fb_naca_points(fb_ctx* ctx, const fb_foil* foil, const fb_pose* pose, int64_t n, const double* points_host,
                   double* sdf_host, double* normals_host) {
  FB_REQUIRE(ctx && foil && pose && points_host && sdf_host, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(n >= 0, FB_ERR_ARG, "negative point count");
  if (n == 0) return FB_OK;
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  double *d_p = nullptr, *d_s = nullptr, *d_n = nullptr;
  FB_HIP(hipMallocAsync((void**)&d_p, n * 2 * sizeof(double), c->stream));
  FB_HIP(hipMallocAsync((void**)&d_s, n * sizeof(double), c->stream));
  if (normals_host) FB_HIP(hipMallocAsync((void**)&d_n, n * 2 * sizeof(double), c->stream));
  FB_HIP(hipMemcpyAsync(d_p, points_host, n * 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(fb::point_sdf_kernel, dim3(blocks), dim3(256), 0, c->stream, fb::make_naca(*foil, *pose), n,
                     d_p, d_s, d_n);
  FB_HIP(hipGetLastError());
  FB_HIP(hipMemcpyAsync(sdf_host, d_s, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (normals_host)
    FB_HIP(hipMemcpyAsync(normals_host, d_n, n * 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FB_HIP(hipFreeAsync(d_p, c->stream));
  FB_HIP(hipFreeAsync(d_s, c->stream));
  if (d_n) FB_HIP(hipFreeAsync(d_n, c->stream));
  FB_HIP(hipStreamSynchronize(c->stream));
  return FB_OK;
}

}  // extern "C"
