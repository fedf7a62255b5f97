// Pressure Poisson solve: preconditioned CG in SciPy's iteration order with the
// reference's DST-I fast-Poisson preconditioner, on gfx950.
//
// Preconditioner (fb/core/_scipy_adapter.py:76-101): z = S (Lambda^-1 o (S r_fluid)) with the
// orthonormal 2D DST-I S, then z = r on solid cells; Jacobi when any axis is periodic.
// Here the y direction is applied as the exact tridiagonal solve it diagonalises,
// (T_y/dy^2 + lambda_x I) z = r_hat (one Thomas sweep per x-mode), and the x direction as a
// DST-I per row computed with a Bluestein chirp-z transform: power-of-two radix-2 FFTs in
// LDS, since the DST-I lengths n+1 (129, 257, 1025, 2049, ...) have large prime factors.
// Both are exact up to rounding, so the CG iterates follow the reference's sequence.
//
// CG (scipy/sparse/linalg/_isolve/iterative.py cg): x0 = 0, r = b; per iteration: stop when
// ||r|| < max(0, rtol ||b||) (strict, tested first); z = M r; rho = r.z;
// p = z (first) or beta p + z; q = A p; alpha = rho / p.q; x += alpha p; r -= alpha q.
// Scalars live on the device; dot products use a fixed-order two-pass reduction so every
// run (and every batch composition) gives the same bits.
#include <cmath>
#include <complex>

#include <atomic>

#include "poisson.h"

namespace fb {

enum : uint8_t { PF_SOLID = 1, PF_RIGHT = 2, PF_UP = 4, PF_WRAPX = 8, PF_WRAPY = 16 };
// Device scalar slots.
enum { S_RN2 = 0, S_RHO, S_RHOPREV, S_PQ, S_ALPHA, S_BETA, S_ATOL, S_DONE, S_ITER, S_FAIL, S_INFO, S_MAXIT, S_DOT,
       S_COUNT = 16,
       // after the two right-hand-side status words at S_COUNT
       S_BSUM = S_COUNT + 2, S_BCNT, S_RES2, S_B2,
       // block ticket of the fused folds (an unsigned int in this double's storage)
       S_TICKET = S_COUNT + 6 };

// Aligned to its full size so that LDS and global accesses move a complex value in one 64-bit
// (FP32) or 128-bit (FP64) instruction.
template <class T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};
template <class T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

struct DstPlan {
  int n = 0, M = 0, logM = 0;
  double scale = 0.0;
  void* chirp = nullptr;  // cplx<T>[n]: exp(-i pi m^2 / (2(n+1))), m = 1..n
  void* fb = nullptr;     // cplx<T>[M]: FFT of the cyclic exp(+i pi m^2 / (2(n+1)))
  void* tw = nullptr;     // cplx<T>[M]: exp(-2 pi i k / M)
};

// ---------------------------------------------------------------- host table set-up

static void host_fft(std::vector<std::complex<long double>>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  const long double pi = 3.141592653589793238462643383279502884L;
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len) {
      for (size_t k = 0; k < len / 2; ++k) {
        long double ang = -2.0L * pi * (long double)k / (long double)len;
        std::complex<long double> w(cosl(ang), sinl(ang));
        auto u = a[i + k];
        auto v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

template <class T>
static int make_plan(DstPlan*& plan, int n, hipStream_t stream) {
  plan = new DstPlan();
  plan->n = n;
  int M = 1, logM = 0;
  while (M < 2 * n - 1) {
    M <<= 1;
    ++logM;
  }
  // the radix-8 Stockham passes give each thread 16 points: at least one full thread per row
  while (M < 16) {
    M <<= 1;
    ++logM;
  }
  if (M > 8192) {
    set_error("DST preconditioner supports at most 4096 cells along x (got %d)", n);
    return FB_ERR_ARG;
  }
  plan->M = M;
  plan->logM = logM;
  plan->scale = std::sqrt(2.0 / (double)(n + 1));
  const long double pi = 3.141592653589793238462643383279502884L;
  const long long fourN = 4LL * (n + 1);
  auto phase = [&](long long m) {  // pi m^2 / (2(n+1)) reduced exactly modulo 2 pi
    long long r = (m * m) % fourN;
    return pi * (long double)r / (long double)(2 * (n + 1));
  };
  std::vector<cplx<T>> chirp(n), fbv(M), tw(M);
  for (int m = 1; m <= n; ++m) {
    long double a = phase(m);
    chirp[m - 1] = {(T)cosl(a), (T)(-sinl(a))};
  }
  std::vector<std::complex<long double>> b(M, {0.0L, 0.0L});
  for (int m = -(n - 1); m <= n - 1; ++m) {
    long double a = phase(m < 0 ? -m : m);
    b[(size_t)((m + M) % M)] = {cosl(a), sinl(a)};
  }
  host_fft(b);
  for (int k = 0; k < M; ++k) fbv[k] = {(T)b[k].real(), (T)b[k].imag()};
  for (int k = 0; k < M; ++k) {
    long double a = -2.0L * pi * (long double)k / (long double)M;
    tw[k] = {(T)cosl(a), (T)sinl(a)};
  }
  FB_HIP(dev_alloc(&plan->chirp, n * sizeof(cplx<T>)));
  FB_HIP(dev_alloc(&plan->fb, M * sizeof(cplx<T>)));
  FB_HIP(dev_alloc(&plan->tw, M * sizeof(cplx<T>)));
  // stream-ordered copies (a legacy-stream copy would wait on every other thread's stream)
  FB_HIP(hipMemcpyAsync(plan->chirp, chirp.data(), n * sizeof(cplx<T>), hipMemcpyHostToDevice, stream));
  FB_HIP(hipMemcpyAsync(plan->fb, fbv.data(), M * sizeof(cplx<T>), hipMemcpyHostToDevice, stream));
  FB_HIP(hipMemcpyAsync(plan->tw, tw.data(), M * sizeof(cplx<T>), hipMemcpyHostToDevice, stream));
  FB_HIP(hipStreamSynchronize(stream));
  return FB_OK;
}

static void free_plan(DstPlan* p) {
  if (!p) return;
  if (p->chirp) (void)dev_free(p->chirp);
  if (p->fb) (void)dev_free(p->fb);
  if (p->tw) (void)dev_free(p->tw);
  delete p;
}

// ---------------------------------------------------------------- device pieces

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  // fixed-order tree: butterfly within each wave, then over waves
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) s += lds[w];
  }
  __syncthreads();
  return s;  // valid in thread 0
}

#define CG_DONE(scal) ((scal)[S_DONE] != 0.0)

// Several grids of one shape can share a launch: blockIdx.z selects the grid, whose vectors sit
// `gs` values apart, its partials `ps` apart and its scalars kScalStride apart.
constexpr int kScalStride = S_COUNT + 8;

// Sum `n` partials in fixed order into scal[slot]; optional post-processing by `mode`. Run by a
// block of 256 threads (a launch of its own, or the last block of the kernel that wrote them).
// mode 0: plain dot -> scal[S_DOT]
// mode 1: rho = sum; beta = rho / rho_prev (iteration > 0)
// mode 2: pq = sum; alpha = rho / pq
// mode 3: rn2 = sum; rho_prev = rho; iteration += 1; then the top-of-loop check
// mode 4: rn2 = sum (initial) and the top-of-loop check
// What the folded sum `s` sets in the grid's scalars, by `mode` (run by one thread).
__device__ __forceinline__ void finalize_tail(double s, double* scal, int mode) {
  if (mode == 0) {
    scal[S_DOT] = s;
    return;
  }
  if (mode == 1) {
    scal[S_RHO] = s;
    if (scal[S_ITER] > 0.0) scal[S_BETA] = s / scal[S_RHOPREV];
    if (!isfinite(s) || !isfinite(scal[S_BETA])) {
      scal[S_FAIL] = 4.0;
      scal[S_DONE] = 1.0;
    }
    return;
  }
  if (mode == 2) {
    scal[S_PQ] = s;
    double a = scal[S_RHO] / s;
    scal[S_ALPHA] = a;
    if (!isfinite(s) || s == 0.0 || !isfinite(a)) {
      scal[S_FAIL] = 4.0;
      scal[S_DONE] = 1.0;
    }
    return;
  }
  if (mode == 3) {
    scal[S_RHOPREV] = scal[S_RHO];
    scal[S_ITER] += 1.0;
  }
  // mode 5 (after mode 4 on ||b||^2, warm start): s = ||b - A x0||^2, the residual tests only
  if (mode == 4) {
    // scipy: atol = max(atol, rtol*||b||); ||b|| == 0 returns b with info 0 before the loop
    scal[S_ATOL] = fmax(0.0, scal[S_ATOL] * sqrt(s));
    if (s == 0.0) {
      scal[S_RN2] = 0.0;
      scal[S_DONE] = 1.0;
      scal[S_INFO] = 0.0;
      return;
    }
  }
  scal[S_RN2] = s;
  if (!isfinite(s)) {
    scal[S_FAIL] = 4.0;
    scal[S_DONE] = 1.0;
    return;
  }
  // top of the next iteration: maxiter exhausted, then the strict residual test
  if (scal[S_ITER] >= scal[S_MAXIT]) {
    scal[S_DONE] = 1.0;
    scal[S_INFO] = scal[S_MAXIT];
  } else if (sqrt(s) < scal[S_ATOL]) {
    scal[S_DONE] = 1.0;
    scal[S_INFO] = 0.0;
  }
}

__device__ __forceinline__ void finalize_body(const double* __restrict__ partial, int n, double* scal, int mode,
                                              double* lds) {
  if (mode != 0 && mode != 4 && CG_DONE(scal)) return;
  // eight independent strided sums per thread, combined pairwise: an order fixed by n alone.
  // Four rounds of eight loads are issued together where n allows, and the tail's loads are
  // unconditional (clamped index, the sum taking them only below n); the additions keep their order.
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int i = threadIdx.x;
  for (; i + 31 * 256 < n; i += 32 * 256) {
    double w[32];
#pragma unroll
    for (int u = 0; u < 32; ++u) w[u] = partial[i + u * 256];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] += w[c * 8 + u];
  }
  for (; i + 7 * 256 < n; i += 8 * 256) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += partial[i + u * 256];
  }
  if (n > 0) {
    double w[7];
#pragma unroll
    for (int u = 0; u < 7; ++u) w[u] = partial[min(i + u * 256, n - 1)];
#pragma unroll
    for (int u = 0; u < 7; ++u)
      if (i + u * 256 < n) a[u] += w[u];
  }
  double s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  s = block_sum<256>(s, lds);
  if (threadIdx.x != 0) return;
  finalize_tail(s, scal, mode);
}


__global__ void __launch_bounds__(256) cg_finalize_kernel(const double* __restrict__ partial, int n, double* scal,
                                                          int mode, int ps) {
  __shared__ double lds[4];
  finalize_body(partial + (size_t)blockIdx.z * ps, n, scal + (size_t)blockIdx.z * kScalStride, mode, lds);
}

// After each block of a 256-thread kernel has written its partial (thread 0), the block that
// finishes last folds all gridDim.x partials with finalize_body: the same sums in the same order
// as a cg_finalize_kernel launch, one launch fewer. The grid's ticket (zeroed by cg_init_kernel)
// counts the finished blocks and is reset by the folding block. Every block takes the ticket with
// one atomic on one address, which serialises, and its release fence writes back the XCD's dirty
// L2 lines; it is used only for grids of at most kFoldBlocks blocks, where the saved launch gap
// outweighs both (with two levels of tickets on every size, 2048 x 1024 FP64 cg_update took 70 us
// against 16 us and a fold launch).
constexpr int kFoldBlocks = 256;
__device__ __forceinline__ void fold_if_last(const double* __restrict__ partial, double* scal, int mode,
                                             double* lds, unsigned nblk) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
    unsigned* ticket = reinterpret_cast<unsigned*>(scal + S_TICKET);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this block's partial before its ticket
    last = atomicAdd(ticket, 1u) == nblk - 1;
    if (last) *ticket = 0u;
  }
  __syncthreads();
  if (!last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every block's partial, not a stale copy
  finalize_body(partial, (int)nblk, scal, mode, lds);
}
__device__ __forceinline__ void fold_if_last(const double* __restrict__ partial, double* scal, int mode,
                                             double* lds) {
  fold_if_last(partial, scal, mode, lds, gridDim.x);
}

template <class T>
__global__ void __launch_bounds__(256) dot_partial_kernel(const T* __restrict__ a, const T* __restrict__ b, size_t n,
                                                          double* __restrict__ partial, int ps) {
  __shared__ double lds[4];
  a += (size_t)blockIdx.z * n;
  b += (size_t)blockIdx.z * n;
  partial += (size_t)blockIdx.z * ps;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  double v = i < n ? (double)a[i] * (double)b[i] : 0.0;
  if constexpr (sizeof(T) == 4) v = i < n ? (double)(a[i] * b[i]) : 0.0;
  v = block_sum<256>(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// Masked 5-point operator row at cell (x, y) applied to a field given by `val` (fb/core/
// _scipy_adapter.py:56-70, same operation order).
template <class T, class F>
__device__ __forceinline__ T apply_row(int x, int y, int nx, int ny, const uint8_t* __restrict__ fl,
                                       const double* __restrict__ diag, T ix2, T iy2, F val) {
  const size_t i = (size_t)y * nx + x;
  const uint8_t f = fl[i];
  T pc = val(x, y);
  if (f & PF_SOLID) return pc;
  T r = (T)diag[i] * pc;
  if (f & PF_RIGHT) r = r - val(x + 1, y) * ix2;
  if (x > 0 && (fl[i - 1] & PF_RIGHT)) r = r - val(x - 1, y) * ix2;
  if (f & PF_UP) r = r - val(x, y + 1) * iy2;
  if (y > 0 && (fl[i - nx] & PF_UP)) r = r - val(x, y - 1) * iy2;
  if (f & PF_WRAPX) {  // fluid pair across the periodic x seam
    if (x == nx - 1) r = r - val(0, y) * ix2;
    if (x == 0) r = r - val(nx - 1, y) * ix2;
  }
  if (f & PF_WRAPY) {
    if (y == ny - 1) r = r - val(x, 0) * iy2;
    if (y == 0) r = r - val(x, ny - 1) * iy2;
  }
  return r;
}

// The masked operator's diagonal at fluid cell (x, y) from the operator flags: setup_kernel's sums
// in its order (fb/core/_scipy_adapter.py:31-54), so the same bits as the stored diagonal. `fl` is
// the cell's flags, `fleft` those of (x-1, y) and `fdown` those of (x, y-1) (0 outside the grid).
__device__ __forceinline__ double diag_from_flags(int x, int y, int nx, int ny, int px, int py, uint8_t fl,
                                                  uint8_t fleft, uint8_t fdown, double ix2, double iy2) {
  double d = 0.0;
  if (fl & PF_RIGHT) d += ix2;
  if (x > 0 && (fleft & PF_RIGHT)) d += ix2;
  if (fl & PF_UP) d += iy2;
  if (y > 0 && (fdown & PF_UP)) d += iy2;
  if (px) {
    if ((fl & PF_WRAPX) && x == nx - 1) d += ix2;
    if ((fl & PF_WRAPX) && x == 0) d += ix2;
  } else {
    if (x == 0) d += ix2;
    if (x == nx - 1) d += ix2;
  }
  if (py) {
    if ((fl & PF_WRAPY) && y == ny - 1) d += iy2;
    if ((fl & PF_WRAPY) && y == 0) d += iy2;
  } else {
    if (y == 0) d += iy2;
    if (y == ny - 1) d += iy2;
  }
  return d;
}

// p_new = first ? z : p_old*beta + z (SciPy: p *= beta; p += z); q = A p_new; partial p.q.
// A block of 256 threads owns a tile of kPqW x TH cells (TH / 4 rows per thread): it forms p once
// over the tile and a one-cell halo in LDS (with the flags of the tile and its left and lower
// neighbours), then applies the operator from LDS with the diagonal recomputed from the flags
// (apply_row's arithmetic and order). Only a periodic seam reads beyond the halo, from global
// memory. Each thread sums its cells' p.q in row order, the block sums its threads, and the tiles'
// partials are summed in tile order (pq_tiles of them per grid). Tall tiles (16 rows) re-read 12%
// of p's inputs as halo instead of 50%; grids with fewer than 1024 such tiles keep 4-row tiles for
// more blocks (measured: 2048 x 1024 FP64 24.3 -> 18.0 us with tall tiles, 1024 x 512 8.7 -> 8.9 us).
constexpr int kPqW = 64;
__host__ __device__ inline int pq_rows(int nx, int ny) {
  return (nx + kPqW - 1) / kPqW * ((ny + 15) / 16) >= 1024 ? 16 : 4;
}
__host__ __device__ inline int pq_tiles(int nx, int ny) {
  return ((nx + kPqW - 1) / kPqW) * ((ny + pq_rows(nx, ny) - 1) / pq_rows(nx, ny));
}

template <class T, int TH>
__global__ void __launch_bounds__(256) cg_pq_kernel(int nx, int ny, const uint8_t* __restrict__ fl,
                                                    const T* __restrict__ z, const T* __restrict__ pold,
                                                    T* __restrict__ pnew, T* __restrict__ q,
                                                    double* __restrict__ partial, double* scal, T ix2, T iy2,
                                                    double ix2d, double iy2d, int px, int py, int ps, int fold) {
  constexpr int RPT = TH / 4;  // rows per thread
  __shared__ T sp[TH + 2][kPqW + 2];
  __shared__ uint8_t sf[TH + 1][kPqW + 1];
  __shared__ double lds[4];
  const size_t N = (size_t)nx * ny;
  {
    const size_t go = (size_t)blockIdx.z * N;
    fl += go;
    z += go;
    pold += go;
    pnew += go;
    q += go;
    partial += (size_t)blockIdx.z * ps;
    scal += (size_t)blockIdx.z * kScalStride;
  }
  // the scalars are read together, ahead of the done test
  const double s_done = scal[S_DONE], s_iter = scal[S_ITER], s_beta = scal[S_BETA];
  if (s_done != 0.0) return;
  const bool first = s_iter == 0.0;
  const T beta = (T)s_beta;
  auto pv = [&](int xx, int yy) -> T {
    const size_t k = (size_t)yy * nx + xx;
    return first ? z[k] : pold[k] * beta + z[k];
  };
  const int x0 = (int)blockIdx.x * kPqW, y0 = (int)blockIdx.y * TH;
  // p over rows y0-1 .. y0+TH and columns x0-1 .. x0+kPqW, flags over rows y0-1 .. y0+TH-1 and
  // columns x0-1 .. x0+kPqW-1. Every load of the thread is issued before its first LDS store, at
  // a clamped index (cell 0) where the cell lies outside the grid, and selected away there.
  constexpr int NPV = (TH + 2) * (kPqW + 2), NFL = (TH + 1) * (kPqW + 1);
  constexpr int UPV = (NPV + 255) / 256, UFL = (NFL + 255) / 256;
  T zv[UPV], ov[UPV];
  uint8_t fv[UFL];
#pragma unroll
  for (int u = 0; u < UPV; ++u) {
    const int k = (int)threadIdx.x + u * 256;
    const int ly = k / (kPqW + 2), lx = k - ly * (kPqW + 2);
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
    const bool in = k < NPV && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    const size_t g = in ? (size_t)gy * nx + gx : 0;
    zv[u] = z[g];
    ov[u] = pold[g];
  }
#pragma unroll
  for (int u = 0; u < UFL; ++u) {
    const int k = (int)threadIdx.x + u * 256;
    const int ly = k / (kPqW + 1), lx = k - ly * (kPqW + 1);
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
    const bool in = k < NFL && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    fv[u] = fl[in ? (size_t)gy * nx + gx : 0];
  }
#pragma unroll
  for (int u = 0; u < UPV; ++u) {
    const int k = (int)threadIdx.x + u * 256;
    const int ly = k / (kPqW + 2), lx = k - ly * (kPqW + 2);
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
    const bool in = gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    if (k < NPV) sp[ly][lx] = in ? (first ? zv[u] : ov[u] * beta + zv[u]) : (T)0;
  }
#pragma unroll
  for (int u = 0; u < UFL; ++u) {
    const int k = (int)threadIdx.x + u * 256;
    const int ly = k / (kPqW + 1), lx = k - ly * (kPqW + 1);
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
    const bool in = gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    if (k < NFL) sf[ly][lx] = in ? fv[u] : (uint8_t)0;
  }
  __syncthreads();
  const int tx = (int)threadIdx.x % kPqW;
  const int x = x0 + tx;
  double v = 0.0;
#pragma unroll
  for (int rr = 0; rr < RPT; ++rr) {
    const int ty = ((int)threadIdx.x / kPqW) * RPT + rr;
    const int y = y0 + ty;
    if (x < nx && y < ny) {
      const size_t i = (size_t)y * nx + x;
      const uint8_t f = sf[ty + 1][tx + 1];
      const T pc = sp[ty + 1][tx + 1];
      T qv = pc;
      if (!(f & PF_SOLID)) {
        const double d = diag_from_flags(x, y, nx, ny, px, py, f, sf[ty + 1][tx], sf[ty][tx + 1], ix2d, iy2d);
        T r = (T)d * pc;
        if (f & PF_RIGHT) r = r - sp[ty + 1][tx + 2] * ix2;
        if (x > 0 && (sf[ty + 1][tx] & PF_RIGHT)) r = r - sp[ty + 1][tx] * ix2;
        if (f & PF_UP) r = r - sp[ty + 2][tx + 1] * iy2;
        if (y > 0 && (sf[ty][tx + 1] & PF_UP)) r = r - sp[ty][tx + 1] * iy2;
        if (f & PF_WRAPX) {  // fluid pair across the periodic x seam
          if (x == nx - 1) r = r - pv(0, y) * ix2;
          if (x == 0) r = r - pv(nx - 1, y) * ix2;
        }
        if (f & PF_WRAPY) {
          if (y == ny - 1) r = r - pv(x, 0) * iy2;
          if (y == 0) r = r - pv(x, ny - 1) * iy2;
        }
        qv = r;
      }
      pnew[i] = pc;
      q[i] = qv;
      v += (sizeof(T) == 4) ? (double)(pc * qv) : (double)pc * (double)qv;
    }
  }
  v = block_sum<256>(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = v;
  if (fold) fold_if_last(partial, scal, 2, lds, gridDim.x * gridDim.y);  // pq and alpha
}

// x += alpha p; r -= alpha q; partial ||r||^2; non-finite iterate flag (the reference's callback).
template <class T>
__global__ void __launch_bounds__(256) cg_update_kernel(size_t N, T* __restrict__ x, T* __restrict__ r,
                                                        const T* __restrict__ p, const T* __restrict__ q,
                                                        double* __restrict__ partial, double* scal, int ps, int fold) {
  __shared__ double lds[4];
  {
    const size_t go = (size_t)blockIdx.z * N;
    x += go;
    r += go;
    p += go;
    q += go;
    partial += (size_t)blockIdx.z * ps;
    scal += (size_t)blockIdx.z * kScalStride;
  }
  const double s_done = scal[S_DONE], s_alpha = scal[S_ALPHA];  // read together
  if (s_done != 0.0) return;
  const T alpha = (T)s_alpha;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  bool bad = false;
  if (i < N) {
    T xv = x[i] + alpha * p[i];
    T rv = r[i] - alpha * q[i];
    x[i] = xv;
    r[i] = rv;
    bad = !isfinite((double)xv);
    v = (sizeof(T) == 4) ? (double)(rv * rv) : (double)rv * (double)rv;
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) scal[S_FAIL] = 3.0;
  v = block_sum<256>(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
  if (fold) fold_if_last(partial, scal, 3, lds);  // ||r||^2 and the top-of-loop check
}

// Jacobi preconditioner (periodic domains): z = (1/max(diag, eps)) r; partial r.z.
template <class T>
__global__ void __launch_bounds__(256) jacobi_kernel(size_t N, const double* __restrict__ diag,
                                                     const T* __restrict__ r, T* __restrict__ z,
                                                     double* __restrict__ partial, const double* scal, double eps) {
  __shared__ double lds[4];
  if (scal && CG_DONE(scal)) return;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (i < N) {
    T inv = (T)(1.0 / fmax(diag[i], eps));
    T zv = inv * r[i];
    z[i] = zv;
    v = (sizeof(T) == 4) ? (double)(r[i] * zv) : (double)r[i] * (double)zv;
  }
  v = block_sum<256>(v, lds);
  if (threadIdx.x == 0 && partial) partial[blockIdx.x] = v;
}

// ---------------------------------------------------------------- DST-I along x

// Radix-R DFT (forward, e^{-2 pi i / R}) of v[0..R) in registers, natural order in and out.
template <class T>
__device__ __forceinline__ void dft2(cplx<T>* v) {
  const cplx<T> a = v[0], b = v[1];
  v[0] = {a.x + b.x, a.y + b.y};
  v[1] = {a.x - b.x, a.y - b.y};
}
template <class T>
__device__ __forceinline__ void dft4(cplx<T>* v) {
  const cplx<T> c0{v[0].x + v[2].x, v[0].y + v[2].y}, c1{v[0].x - v[2].x, v[0].y - v[2].y};
  const cplx<T> c2{v[1].x + v[3].x, v[1].y + v[3].y};
  const cplx<T> c3{v[1].y - v[3].y, v[3].x - v[1].x};  // (v1 - v3) * (-i)
  v[0] = {c0.x + c2.x, c0.y + c2.y};
  v[2] = {c0.x - c2.x, c0.y - c2.y};
  v[1] = {c1.x + c3.x, c1.y + c3.y};
  v[3] = {c1.x - c3.x, c1.y - c3.y};
}
template <class T>
__device__ __forceinline__ void dft8(cplx<T>* v) {
  const T h = (T)0.70710678118654752440084436210485;
  // first radix-2 stage (pairs s, s+4), odd half twiddled by e^{-2 pi i s / 8}
  cplx<T> e[4], o[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    e[s] = {v[s].x + v[s + 4].x, v[s].y + v[s + 4].y};
    o[s] = {v[s].x - v[s + 4].x, v[s].y - v[s + 4].y};
  }
  o[1] = {(o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h};
  o[2] = {o[2].y, -o[2].x};
  o[3] = {(o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h};
  dft4(e);
  dft4(o);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    v[2 * m] = e[m];
    v[2 * m + 1] = o[m];
  }
}
template <class T, int R>
__device__ __forceinline__ void dft_r(cplx<T>* v) {
  if constexpr (R == 8) dft8(v);
  else if constexpr (R == 4) dft4(v);
  else dft2(v);
}

// LDS slot of point i of a row: one pad slot after every 8 points, so that the writes of the
// first Stockham passes (lanes 8 or 64 points apart) land in different banks. A row of M points
// takes lds_row(M) slots.
__host__ __device__ constexpr int lds_slot(int i) { return i + (i >> 3); }
__host__ __device__ constexpr int lds_row(int M) { return M + (M >> 3); }

// One radix-R Stockham pass (sub-transforms of length L -> L*R) over a row of M points: each of
// the row's M/PPT threads takes its PPT points through `ld` (the row in LDS, or the pass's input
// straight from global memory), the block synchronises when the pass reads and writes the same
// LDS row (SYNC_IN), and the results go through `st` to their autosorted places (the LDS row, or
// the transform's consumer).
template <class T, int R, int PPT, bool SYNC_IN, class Ld, class St>
__device__ __forceinline__ void stockham_pass(Ld&& ld, St&& st, const cplx<T>* __restrict__ roots, int M, int L,
                                              int t, int nthr) {
  constexpr int NB = PPT / R;
  const int MR = M / R;
  cplx<T> v[NB][R];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int b = t + u * nthr;
#pragma unroll
    for (int s = 0; s < R; ++s) v[u][s] = ld(b + s * MR);
  }
  if constexpr (SYNC_IN) __syncthreads();
  const int stride = M / (L * R);
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int b = t + u * nthr;
    const int k = b & (L - 1);
    // one table load per butterfly; the higher twiddles are its powers (table loads of every
    // power were measured slower: the lanes' addresses scatter). The first pass (L = 1) has none.
    if (L > 1) {
      const cplx<T> w1 = roots[k * stride];
      cplx<T> ws = w1;
#pragma unroll
      for (int s = 1; s < R; ++s) {
        v[u][s] = cmul(v[u][s], ws);
        if (s + 1 < R) ws = cmul(ws, w1);
      }
    }
    dft_r<T, R>(v[u]);
    const int base = (b - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) st(base + q * L, v[u][q], u * R + q);
  }
  __syncthreads();
}

// Forward FFT of each row's M = 2^logM points (M >= 16: radix-8 passes, then one radix-2 or -4
// pass), with M / PPT threads per row, each taking PPT points of every pass. The first pass reads
// through `first_ld` (SYNC_FIRST when that reads the LDS row), the last pass writes through
// `last_st`; the passes between keep the row in LDS.
template <class T, int PPT, bool SYNC_FIRST, class Ld, class St>
__device__ __forceinline__ void fft_rows(cplx<T>* buf, const cplx<T>* __restrict__ roots, int M, int logM, int t,
                                         int nthr, Ld&& first_ld, St&& last_st) {
  auto lds_ld = [buf](int i) { return buf[lds_slot(i)]; };
  auto lds_st = [buf](int i, cplx<T> v, int) { buf[lds_slot(i)] = v; };
  const int n8 = logM / 3, rem = logM % 3;
  const int last8 = rem ? -1 : n8 - 1;  // the last radix-8 pass is the transform's last pass
  int L = 1;
  for (int p = 0; p < n8; ++p, L *= 8) {
    if (p == 0) stockham_pass<T, 8, PPT, SYNC_FIRST>(first_ld, lds_st, roots, M, L, t, nthr);
    else if (p == last8) stockham_pass<T, 8, PPT, true>(lds_ld, last_st, roots, M, L, t, nthr);
    else stockham_pass<T, 8, PPT, true>(lds_ld, lds_st, roots, M, L, t, nthr);
  }
  if (rem == 2) stockham_pass<T, 4, PPT, true>(lds_ld, last_st, roots, M, L, t, nthr);
  else if (rem == 1) stockham_pass<T, 2, PPT, true>(lds_ld, last_st, roots, M, L, t, nthr);
}

// Orthonormal DST-I along x of every row, Bluestein chirp-z with Stockham FFTs in LDS. A block
// holds `rows` rows of M points, M/PPT threads each (PPT 16 or 8: more threads per row for short
// grids; the results are the same bits either way).
// The transform runs in T; `in` (TI) and `out`, `rsrc` (TO) may be wider (an FP64 residual through
// the FP32 transform). The product with the transformed chirp kernel is applied as the first
// FFT's last pass stores. FUSE bit 1: the chirp product feeds the first FFT pass straight from
// global memory; bit 2: the second FFT's last pass hands its points to the output. Without them
// (fewer registers) the input and output go through the LDS row. The arithmetic is the same.
// MODE 0: input r masked to fluid cells, output to `out`.
// MODE 1: output z on fluid cells, z = r on solid cells, partial r.z per row, summed in an order
// fixed by M alone: M/16 strided sums, then a pairwise tree.
// MODE is a template parameter so that neither stage branches per point: every point's loads
// (at a clamped index where the point lies past the row) are issued together, and only the
// stores are predicated.
template <class T, class TI, class TO, int PPT, int FUSE, int MODE>
__global__ void __launch_bounds__(512) dst_rows_kernel(int nx, int ny, const TI* __restrict__ in,
                                                       const uint8_t* __restrict__ fl, const TO* __restrict__ rsrc,
                                                       TO* __restrict__ out, const cplx<T>* __restrict__ chirp,
                                                       const cplx<T>* __restrict__ fbt,
                                                       const cplx<T>* __restrict__ roots, int M, int logM, int rows,
                                                       T scale, double* __restrict__ partial,
                                                       const double* scal, int ps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  {
    const size_t go = (size_t)blockIdx.z * nx * ny;
    in += go;
    fl += go;
    rsrc += go;
    out += go;
    partial += (size_t)blockIdx.z * ps;
    if (scal) scal += (size_t)blockIdx.z * kScalStride;
  }
  if (scal && CG_DONE(scal)) return;
  const int nthr = M / PPT;
  const int nred = M >> 4;
  const int lr = threadIdx.x / nthr;  // row within the block
  const int t = threadIdx.x - lr * nthr;
  const int row = blockIdx.x * rows + lr;
  const bool live = row < ny;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw) + (size_t)lr * lds_row(M);
  double* red = reinterpret_cast<double*>(smem_raw + (size_t)rows * lds_row(M) * sizeof(cplx<T>)) + (size_t)lr * nred;
  double* prod = reinterpret_cast<double*>(buf);  // the row's r.z products, once the transform is done
  const size_t base = (size_t)(live ? row : 0) * nx;
  // first FFT: the row times the chirp (zero past nx), then the convolution kernel's spectrum,
  // conjugated for the inverse FFT taken as conj(FFT(conj(.)))
  // The loads take a clamped index (base is a valid row, column 0 stays inside it) and are
  // issued whatever the point's fate, so that the points of a pass have their loads in flight
  // together rather than one branch and one memory round trip after another.
  auto chirped = [&](int j) -> cplx<T> {
    const bool ok = live && j < nx;
    const size_t jj = ok ? (size_t)j : 0;
    T v = (T)in[base + jj];
    const uint8_t f = fl[base + jj];
    const cplx<T> c = chirp[jj];
    if (MODE == 0 && (f & PF_SOLID)) v = (T)0;
    return ok ? cplx<T>{v * c.x, v * c.y} : cplx<T>{(T)0, (T)0};
  };
  auto spectrum = [&](int k, cplx<T> w, int) {
    const cplx<T> v = cmul(w, fbt[k]);
    buf[lds_slot(k)] = {v.x, -v.y};
  };
  auto lds_ld = [buf](int i) { return buf[lds_slot(i)]; };
  if constexpr ((FUSE & 1) != 0) {
    fft_rows<T, PPT, false>(buf, roots, M, logM, t, nthr, chirped, spectrum);
  } else {
    for (int j = t; j < M; j += nthr) buf[lds_slot(j)] = chirped(j);
    __syncthreads();
    fft_rows<T, PPT, true>(buf, roots, M, logM, t, nthr, lds_ld, spectrum);
  }
  // second FFT: its outputs k < nx times the chirp, scaled, are the transform (loads at a
  // clamped index, only the stores predicated)
  const T invM = (T)1 / (T)M;
  auto output = [&](int k, cplx<T> w, int) {
    const bool ok = live && k < nx;
    const size_t kk = ok ? (size_t)k : 0;
    const cplx<T> ck = chirp[kk];
    const cplx<T> conv{w.x * invM, -w.y * invM};
    const cplx<T> y = cmul(ck, conv);
    const T dv = -y.y * scale;
    if constexpr (MODE == 0) {
      if (ok) out[base + k] = (TO)dv;
    } else {
      const TO rv = rsrc[base + kk];
      const TO zv = (fl[base + kk] & PF_SOLID) ? rv : (TO)dv;
      if (ok) {
        out[base + k] = zv;
        prod[k] = (sizeof(TO) == 4) ? (double)(rv * zv) : (double)rv * (double)zv;
      }
    }
  };
  if constexpr ((FUSE & 2) != 0) {
    fft_rows<T, PPT, true>(buf, roots, M, logM, t, nthr, lds_ld, output);
  } else {
    auto lds_st = [buf](int i, cplx<T> v, int) { buf[lds_slot(i)] = v; };
    fft_rows<T, PPT, true>(buf, roots, M, logM, t, nthr, lds_ld, lds_st);
    // this thread's outputs k = t + j nthr < nx (nx <= M / 2, so at most PPT / 2 of them), read
    // before any product overwrites the row
    constexpr int KMAX = PPT / 2;
    cplx<T> w[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      const int k = t + j * nthr;
      w[j] = k < M ? buf[lds_slot(k)] : cplx<T>{(T)0, (T)0};
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < KMAX; ++j) output(t + j * nthr, w[j], j);
    __syncthreads();
  }
  if constexpr (MODE == 0) return;
  // the products are summed in the fixed order (the last writes are behind a barrier)
  for (int v = t; v < nred; v += nthr) {
    double acc = 0.0;
    if (live)
      for (int k = v; k < nx; k += nred) acc += prod[k];
    red[v] = acc;
  }
  __syncthreads();
  for (int s2 = nred >> 1; s2 > 0; s2 >>= 1) {
    for (int v = t; v < s2; v += nthr) red[v] += red[v + s2];
    __syncthreads();
  }
  if (t == 0 && live) partial[row] = red[0];
}

// ---------------------------------------------------------------- tridiagonal solves along y

// Partition solve of (T_y/dy^2 + ex[i] I) z = w for every x-mode column i, in place. The ny rows
// are cut into P chunks of L rows (the last has Llast), and one launch runs three stages over a
// block of 16 modes:
//  1. each chunk of each mode is eliminated with the chunk's LU factors and zero boundary values,
//     keeping the chunk's first and last value (one thread per chunk, blockDim.y chunks at a time);
//  2. the 2P interface values of each mode: two affine recurrences over the chunks, each a
//     segmented scan over blockDim.y segments of ceil(P/blockDim.y) chunks;
//  3. each chunk adds its two spikes scaled by the neighbouring interface values.
// blockDim.y (16 or 32) depends on P alone (tri_threads), so every grid and batch size gives the
// same bits.
// Tables are [chunk type][quantity][row][mode] so that 16 neighbouring modes read 16 neighbours.
constexpr int kTriL = 8;  // rows per chunk
constexpr int kTriFusedMaxP = 32;  // chunks per column up to which the three stages run as one launch
// Chunk threads (and scan segments) per mode of the one-launch form: one per chunk, at least 16.
inline int tri_threads(int P) { return P <= 16 ? 16 : 32; }

template <class T>
__global__ void __launch_bounds__(512) tri_solve_kernel(int nx, int ny, T* __restrict__ w, const T* __restrict__ fac,
                                                         int Llast, int P, T a, const T* __restrict__ itf,
                                                         T* __restrict__ yF, T* __restrict__ yE, const double* scal) {
  w += (size_t)blockIdx.z * nx * ny;
  yF += (size_t)blockIdx.z * P * nx;
  yE += (size_t)blockIdx.z * P * nx;
  if (scal) scal += (size_t)blockIdx.z * kScalStride;
  if (scal && CG_DONE(scal)) return;
  __shared__ T sA[32][16], sB[32][16], sS[32][16];
  const int m = threadIdx.x, ct = threadIdx.y;
  const int NS = (int)blockDim.y;
  const int i = blockIdx.x * 16 + m;
  const bool on = i < nx;
  // 1. chunk elimination, forward then back substitution within the chunk
  if (on)
    for (int c = ct; c < P; c += (int)blockDim.y) {
      const int last = c == P - 1;
      const int Lc = last ? Llast : kTriL;
      const T* invm = fac + (size_t)(last * 4 + 0) * kTriL * nx + i;
      const T* cp = fac + (size_t)(last * 4 + 1) * kTriL * nx + i;
      T* col = w + (size_t)c * kTriL * nx + i;
      // every load first (rows past a short last chunk read its last row; the factor tables hold
      // kTriL rows for either chunk type), so that they are in flight together
      T v[kTriL], fi[kTriL], fc[kTriL];
#pragma unroll
      for (int j = 0; j < kTriL; ++j) {
        v[j] = col[(size_t)(j < Lc ? j : Lc - 1) * nx];
        fi[j] = invm[(size_t)j * nx];
        fc[j] = cp[(size_t)j * nx];
      }
      T y = (T)0;
#pragma unroll
      for (int j = 0; j < kTriL; ++j)
        if (j < Lc) {
          y = (v[j] - a * y) * fi[j];
          v[j] = y;
        }
      T yn = y;
#pragma unroll
      for (int j = kTriL - 2; j >= 0; --j)
        if (j < Lc - 1) {
          yn = v[j] - fc[j] * yn;
          v[j] = yn;
        }
#pragma unroll
      for (int j = 0; j < kTriL; ++j)
        if (j < Lc) col[(size_t)j * nx] = v[j];
      yF[(size_t)c * nx + i] = v[0];
      yE[(size_t)c * nx + i] = y;
    }
  __syncthreads();
  // 2. interface recurrences. Forward: r_k = g yF_k + h pp_{k-1}, pp_k = A pp_{k-1} + (yE_k + mg yF_k);
  // back: F_k = sv F_{k+1} + r_k, E_k = pp_k + qq F_{k+1} (F_P = 0). Each is an affine recurrence:
  // every thread composes its segment's maps, one thread per mode chains the NS segment maps, then
  // each thread replays its segment from its start value. itf = [g, h, A, mg, sv, qq][chunk][mode].
  const int gseg = ct;
  const int S = (P + NS - 1) / NS;
  const int k0 = min(P, gseg * S), k1 = min(P, k0 + S);
  const size_t PQ = (size_t)P * nx;
  const T* cg = itf;
  const T* ch = itf + PQ;
  const T* cA = itf + 2 * PQ;
  const T* cmg = itf + 3 * PQ;
  const T* csv = itf + 4 * PQ;
  const T* cqq = itf + 5 * PQ;
  T Ac = (T)1, Bc = (T)0;
  if (on)
    for (int k = k0; k < k1; ++k) {
      const size_t o = (size_t)k * nx + i;
      const T A = cA[o];
      Ac = A * Ac;
      Bc = A * Bc + (yE[o] + cmg[o] * yF[o]);
    }
  sA[gseg][m] = Ac;
  sB[gseg][m] = Bc;
  __syncthreads();
  if (gseg == 0) {
    T cur = (T)0;
    for (int q = 0; q < NS; ++q) {
      sS[q][m] = cur;
      cur = sA[q][m] * cur + sB[q][m];
    }
  }
  __syncthreads();
  if (on) {
    T pp = sS[gseg][m];
    for (int k = k0; k < k1; ++k) {
      const size_t o = (size_t)k * nx + i;
      const T f = yF[o];
      const T r = cg[o] * f + ch[o] * pp;
      pp = cA[o] * pp + (yE[o] + cmg[o] * f);
      yF[o] = r;
      yE[o] = pp;
    }
  }
  Ac = (T)1;
  Bc = (T)0;
  if (on)
    for (int k = k1 - 1; k >= k0; --k) {
      const size_t o = (size_t)k * nx + i;
      const T sv = csv[o];
      Ac = sv * Ac;
      Bc = sv * Bc + yF[o];
    }
  __syncthreads();
  sA[gseg][m] = Ac;
  sB[gseg][m] = Bc;
  __syncthreads();
  if (gseg == 0) {
    T cur = (T)0;
    for (int q = NS - 1; q >= 0; --q) {
      sS[q][m] = cur;
      cur = sA[q][m] * cur + sB[q][m];
    }
  }
  __syncthreads();
  if (on) {
    T Fn = sS[gseg][m];
    for (int k = k1 - 1; k >= k0; --k) {
      const size_t o = (size_t)k * nx + i;
      const T F = yF[o] + csv[o] * Fn;
      const T E = yE[o] + cqq[o] * Fn;
      yF[o] = F;
      yE[o] = E;
      Fn = F;
    }
  }
  // the neighbouring segments' interface values, written by other threads of this block
  __syncthreads();
  // 3. spikes
  if (on)
    for (int c = ct; c < P; c += (int)blockDim.y) {
      const int last = c == P - 1;
      const int Lc = last ? Llast : kTriL;
      const T* al = fac + (size_t)(last * 4 + 2) * kTriL * nx + i;
      const T* be = fac + (size_t)(last * 4 + 3) * kTriL * nx + i;
      const T El = c > 0 ? yE[(size_t)(c - 1) * nx + i] : (T)0;
      const T Fr = c < P - 1 ? yF[(size_t)(c + 1) * nx + i] : (T)0;
      T* col = w + (size_t)c * kTriL * nx + i;
      T cv[kTriL], av[kTriL], bv[kTriL];
#pragma unroll
      for (int j = 0; j < kTriL; ++j) {
        cv[j] = col[(size_t)(j < Lc ? j : Lc - 1) * nx];
        av[j] = al[(size_t)j * nx];
        bv[j] = be[(size_t)j * nx];
      }
#pragma unroll
      for (int j = 0; j < kTriL; ++j)
        if (j < Lc) col[(size_t)j * nx] = cv[j] + av[j] * El + bv[j] * Fr;
    }
}

// The same three stages as three launches, for long columns (P > kTriFusedMaxP): a thread per mode
// and chunk in the first and last (64 neighbouring modes to a wave), a 16 x 16 block of modes and
// scan segments in the second. Measured faster than the one-launch form there (1024 x 512 FP64
// preconditioner 50.5 against 52.1 us; 2048 x 1024 117.4 against 119.4 us), and slower below
// (512 x 256: 37.6 against 35.8 us).
template <class T>
__global__ void __launch_bounds__(256) tri_chunk_kernel(int nx, int ny, T* __restrict__ w,
                                                        const T* __restrict__ fac, int Llast, int P, T a,
                                                        T* __restrict__ yF, T* __restrict__ yE, const double* scal) {
  w += (size_t)blockIdx.z * nx * ny;
  yF += (size_t)blockIdx.z * P * nx;
  yE += (size_t)blockIdx.z * P * nx;
  if (scal) scal += (size_t)blockIdx.z * kScalStride;
  if (scal && CG_DONE(scal)) return;
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int c = blockIdx.y * 4 + threadIdx.y;
  if (i >= nx || c >= P) return;
  const int last = c == P - 1;
  const int Lc = last ? Llast : kTriL;
  const T* invm = fac + (size_t)(last * 4 + 0) * kTriL * nx + i;
  const T* cp = fac + (size_t)(last * 4 + 1) * kTriL * nx + i;
  T* col = w + (size_t)c * kTriL * nx + i;
  // every load first, as in tri_solve_kernel
  T v[kTriL], fi[kTriL], fc[kTriL];
#pragma unroll
  for (int j = 0; j < kTriL; ++j) {
    v[j] = col[(size_t)(j < Lc ? j : Lc - 1) * nx];
    fi[j] = invm[(size_t)j * nx];
    fc[j] = cp[(size_t)j * nx];
  }
  // forward elimination, then back substitution, within the chunk
  T y = (T)0;
#pragma unroll
  for (int j = 0; j < kTriL; ++j)
    if (j < Lc) {
      y = (v[j] - a * y) * fi[j];
      v[j] = y;
    }
  T yn = y;
#pragma unroll
  for (int j = kTriL - 2; j >= 0; --j)
    if (j < Lc - 1) {
      yn = v[j] - fc[j] * yn;
      v[j] = yn;
    }
#pragma unroll
  for (int j = 0; j < kTriL; ++j)
    if (j < Lc) col[(size_t)j * nx] = v[j];
  yF[(size_t)c * nx + i] = v[0];
  yE[(size_t)c * nx + i] = y;
}

// Interface recurrences per mode, as segmented scans: 16 modes x 16 segments of chunks per
// block. Forward: r_k = g yF_k + h pp_{k-1}, pp_k = A pp_{k-1} + (yE_k + mg yF_k); back:
// F_k = sv F_{k+1} + r_k, E_k = pp_k + qq F_{k+1} (F_P = 0). Each is an affine recurrence: every
// thread composes its segment's maps, one thread per mode chains the 16 segment maps, then each
// thread replays its segment from its start value. itf = [g, h, A, mg, sv, qq][chunk][mode].
template <class T, int SMAX>
__global__ void __launch_bounds__(256) tri_interface_kernel(int nx, int P, const T* __restrict__ itf,
                                                            T* __restrict__ yF, T* __restrict__ yE,
                                                            const double* scal) {
  yF += (size_t)blockIdx.z * P * nx;
  yE += (size_t)blockIdx.z * P * nx;
  if (scal) scal += (size_t)blockIdx.z * kScalStride;
  if (scal && CG_DONE(scal)) return;
  __shared__ T sA[16][16], sB[16][16], sS[16][16];
  const int m = threadIdx.x, gseg = threadIdx.y;
  const int i = blockIdx.x * 16 + m;
  const bool on = i < nx;
  const int S = (P + 15) / 16;
  const int k0 = min(P, gseg * S), k1 = min(P, k0 + S);
  const size_t PQ = (size_t)P * nx;
  const T* cg = itf;
  const T* ch = itf + PQ;
  const T* cA = itf + 2 * PQ;
  const T* cmg = itf + 3 * PQ;
  const T* csv = itf + 4 * PQ;
  const T* cqq = itf + 5 * PQ;
  if constexpr (SMAX > 0) {
    // the segment (S <= SMAX chunks) is loaded once into registers, every load in flight
    // together, and the four passes below run on it in the order of the loop form
    T vF[SMAX], vE[SMAX], vg[SMAX], vh[SMAX], vA[SMAX], vmg[SMAX], vsv[SMAX], vqq[SMAX];
#pragma unroll
    for (int j = 0; j < SMAX; ++j) {
      const bool ok = on && k0 + j < k1;
      const size_t o = ok ? (size_t)(k0 + j) * nx + i : 0;
      vF[j] = yF[o];
      vE[j] = yE[o];
      vg[j] = cg[o];
      vh[j] = ch[o];
      vA[j] = cA[o];
      vmg[j] = cmg[o];
      vsv[j] = csv[o];
      vqq[j] = cqq[o];
    }
    T Ac = (T)1, Bc = (T)0;
#pragma unroll
    for (int j = 0; j < SMAX; ++j)
      if (k0 + j < k1) {
        Ac = vA[j] * Ac;
        Bc = vA[j] * Bc + (vE[j] + vmg[j] * vF[j]);
      }
    sA[gseg][m] = Ac;
    sB[gseg][m] = Bc;
    __syncthreads();
    if (gseg == 0) {
      T cur = (T)0;
      for (int q = 0; q < 16; ++q) {
        sS[q][m] = cur;
        cur = sA[q][m] * cur + sB[q][m];
      }
    }
    __syncthreads();
    T pp = sS[gseg][m];
#pragma unroll
    for (int j = 0; j < SMAX; ++j)
      if (k0 + j < k1) {
        const T f = vF[j];
        const T r = vg[j] * f + vh[j] * pp;
        pp = vA[j] * pp + (vE[j] + vmg[j] * f);
        vF[j] = r;
        vE[j] = pp;
      }
    Ac = (T)1;
    Bc = (T)0;
#pragma unroll
    for (int j = SMAX - 1; j >= 0; --j)
      if (k0 + j < k1) {
        Ac = vsv[j] * Ac;
        Bc = vsv[j] * Bc + vF[j];
      }
    __syncthreads();
    sA[gseg][m] = Ac;
    sB[gseg][m] = Bc;
    __syncthreads();
    if (gseg == 0) {
      T cur = (T)0;
      for (int q = 15; q >= 0; --q) {
        sS[q][m] = cur;
        cur = sA[q][m] * cur + sB[q][m];
      }
    }
    __syncthreads();
    T Fn = sS[gseg][m];
#pragma unroll
    for (int j = SMAX - 1; j >= 0; --j)
      if (k0 + j < k1) {
        const T F = vF[j] + vsv[j] * Fn;
        const T E = vE[j] + vqq[j] * Fn;
        vF[j] = F;
        vE[j] = E;
        Fn = F;
      }
    if (on)
#pragma unroll
      for (int j = 0; j < SMAX; ++j)
        if (k0 + j < k1) {
          const size_t o = (size_t)(k0 + j) * nx + i;
          yF[o] = vF[j];
          yE[o] = vE[j];
        }
    return;
  } else {
    // forward: compose pp_{k1-1} = Ac pp_{k0-1} + Bc
    T Ac = (T)1, Bc = (T)0;
    if (on)
      for (int k = k0; k < k1; ++k) {
        const size_t o = (size_t)k * nx + i;
        const T A = cA[o];
        Ac = A * Ac;
        Bc = A * Bc + (yE[o] + cmg[o] * yF[o]);
      }
    sA[gseg][m] = Ac;
    sB[gseg][m] = Bc;
    __syncthreads();
    if (gseg == 0) {
      T cur = (T)0;
      for (int q = 0; q < 16; ++q) {
        sS[q][m] = cur;
        cur = sA[q][m] * cur + sB[q][m];
      }
    }
    __syncthreads();
    if (on) {
      T pp = sS[gseg][m];
      for (int k = k0; k < k1; ++k) {
        const size_t o = (size_t)k * nx + i;
        const T f = yF[o];
        const T r = cg[o] * f + ch[o] * pp;
        pp = cA[o] * pp + (yE[o] + cmg[o] * f);
        yF[o] = r;
        yE[o] = pp;
      }
    }
    // back: compose F_{k0} = Ac F_{k1} + Bc
    Ac = (T)1;
    Bc = (T)0;
    if (on)
      for (int k = k1 - 1; k >= k0; --k) {
        const size_t o = (size_t)k * nx + i;
        const T sv = csv[o];
        Ac = sv * Ac;
        Bc = sv * Bc + yF[o];
      }
    __syncthreads();
    sA[gseg][m] = Ac;
    sB[gseg][m] = Bc;
    __syncthreads();
    if (gseg == 0) {
      T cur = (T)0;
      for (int q = 15; q >= 0; --q) {
        sS[q][m] = cur;
        cur = sA[q][m] * cur + sB[q][m];
      }
    }
    __syncthreads();
    if (on) {
      T Fn = sS[gseg][m];
      for (int k = k1 - 1; k >= k0; --k) {
        const size_t o = (size_t)k * nx + i;
        const T F = yF[o] + csv[o] * Fn;
        const T E = yE[o] + cqq[o] * Fn;
        yF[o] = F;
        yE[o] = E;
        Fn = F;
      }
    }

  }
}

template <class T>
__global__ void __launch_bounds__(256) tri_spike_kernel(int nx, int ny, T* __restrict__ w, const T* __restrict__ fac,
                                                        int Llast, int P, const T* __restrict__ F,
                                                        const T* __restrict__ E, const double* scal) {
  w += (size_t)blockIdx.z * nx * ny;
  F += (size_t)blockIdx.z * P * nx;
  E += (size_t)blockIdx.z * P * nx;
  if (scal) scal += (size_t)blockIdx.z * kScalStride;
  if (scal && CG_DONE(scal)) return;
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int c = blockIdx.y * 4 + threadIdx.y;
  if (i >= nx || c >= P) return;
  const int last = c == P - 1;
  const int Lc = last ? Llast : kTriL;
  const T* al = fac + (size_t)(last * 4 + 2) * kTriL * nx + i;
  const T* be = fac + (size_t)(last * 4 + 3) * kTriL * nx + i;
  const T El = c > 0 ? E[(size_t)(c - 1) * nx + i] : (T)0;
  const T Fr = c < P - 1 ? F[(size_t)(c + 1) * nx + i] : (T)0;
  T* col = w + (size_t)c * kTriL * nx + i;
  T cv[kTriL], av[kTriL], bv[kTriL];
#pragma unroll
  for (int j = 0; j < kTriL; ++j) {
    cv[j] = col[(size_t)(j < Lc ? j : Lc - 1) * nx];
    av[j] = al[(size_t)j * nx];
    bv[j] = be[(size_t)j * nx];
  }
#pragma unroll
  for (int j = 0; j < kTriL; ++j)
    if (j < Lc) col[(size_t)j * nx] = cv[j] + av[j] * El + bv[j] * Fr;
}

template <class T>
__global__ void __launch_bounds__(256) copy_kernel(size_t N, const T* __restrict__ a, T* __restrict__ b,
                                                   const double* scal) {
  if (scal && CG_DONE(scal)) return;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N) b[i] = a[i];
}

__global__ void __launch_bounds__(256) setup_kernel(int nx, int ny, const uint8_t* __restrict__ solid, int px, int py,
                                                    double ix2, double iy2, uint8_t* __restrict__ fl,
                                                    double* __restrict__ diag) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  int x = (int)(i % nx), y = (int)(i / nx);
  auto fluid = [&](int xx, int yy) { return solid[(size_t)yy * nx + xx] == 0; };
  bool f = fluid(x, y);
  uint8_t b = f ? 0 : PF_SOLID;
  if (f && x + 1 < nx && fluid(x + 1, y)) b |= PF_RIGHT;
  if (f && y + 1 < ny && fluid(x, y + 1)) b |= PF_UP;
  if (px && f && ((x == nx - 1 && fluid(0, y)) || (x == 0 && fluid(nx - 1, y)))) b |= PF_WRAPX;
  if (py && f && ((y == ny - 1 && fluid(x, 0)) || (y == 0 && fluid(x, ny - 1)))) b |= PF_WRAPY;
  fl[i] = b;
  // diagonal in the reference's accumulation order (fb/core/_scipy_adapter.py:31-54)
  double d = 0.0;
  if (x + 1 < nx && f && fluid(x + 1, y)) d += ix2;
  if (x > 0 && f && fluid(x - 1, y)) d += ix2;
  if (y + 1 < ny && f && fluid(x, y + 1)) d += iy2;
  if (y > 0 && f && fluid(x, y - 1)) d += iy2;
  if (px) {
    if (x == nx - 1 && f && fluid(0, y)) d += ix2;
    if (x == 0 && f && fluid(nx - 1, y)) d += ix2;
  } else {
    if (x == 0 && f) d += ix2;
    if (x == nx - 1 && f) d += ix2;
  }
  if (py) {
    if (y == ny - 1 && f && fluid(x, 0)) d += iy2;
    if (y == 0 && f && fluid(x, ny - 1)) d += iy2;
  } else {
    if (y == 0 && f) d += iy2;
    if (y == ny - 1 && f) d += iy2;
  }
  diag[i] = f ? d : 1.0;
}

// q = A p in FP64 (true residual); r = b - q with partial ||r||^2 into partial.
__global__ void __launch_bounds__(256) residual_kernel(int nx, int ny, const uint8_t* __restrict__ fl,
                                                       const double* __restrict__ diag, const double* __restrict__ b,
                                                       const double* __restrict__ x, double* __restrict__ r,
                                                       double* __restrict__ partial, double ix2, double iy2, int ps) {
  __shared__ double lds[4];
  {
    const size_t go = (size_t)blockIdx.z * nx * ny;
    fl += go;
    diag += go;
    b += go;
    x += go;
    r += go;
    partial += (size_t)blockIdx.z * ps;
  }
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (i < (size_t)nx * ny) {
    int xx = (int)(i % nx), yy = (int)(i / nx);
    auto xv = [&](int a, int c) -> double { return x[(size_t)c * nx + a]; };
    double ax = apply_row<double>(xx, yy, nx, ny, fl, diag, ix2, iy2, xv);
    double rv = b[i] - ax;
    r[i] = rv;
    v = rv * rv;
  }
  v = block_sum<256>(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// RHS preparation (fb/core/_scipy_adapter.py:108-120): b = rhs on fluid, 0 on solid;
// block partials of the fluid sum and count; non-finite count and max |b| via order-free atomics.
__global__ void __launch_bounds__(256) rhs_prep_kernel(size_t N, const double* __restrict__ rhs,
                                                       const uint8_t* __restrict__ fl, double* __restrict__ b,
                                                       double* __restrict__ partial, double* __restrict__ partial_n,
                                                       unsigned long long* __restrict__ stat, int ps) {
  __shared__ double lds[4];
  {
    const size_t go = (size_t)blockIdx.z * N;
    rhs += go;
    fl += go;
    b += go;
    partial += (size_t)blockIdx.z * ps;
    partial_n += (size_t)blockIdx.z * ps;
    stat += (size_t)blockIdx.z * kScalStride;  // the slots sit in the grid's scalar block
  }
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  double v = 0.0, c = 0.0, m = 0.0;
  bool nf = false;
  if (i < N) {
    bool fluid = !(fl[i] & PF_SOLID);
    double x = fluid ? rhs[i] : 0.0;
    b[i] = x;
    if (fluid) {
      v = x;
      c = 1.0;
    }
    nf = !isfinite(x);
    m = nf ? 0.0 : fabs(x);
  }
  if (__any(nf) && (threadIdx.x & 63) == 0) atomicAdd(&stat[0], 1ull);
  // the largest |b| of the block (the bits of a non-negative double order like the value), then one
  // maximum per block into the grid's slot: every wave's requests on that one address serialise
  __shared__ unsigned long long lmax[4];
  const unsigned long long mk = block_max_u64((unsigned long long)__double_as_longlong(m), lmax);
  if (threadIdx.x == 0) key_max(&stat[1], mk);
  double sv = block_sum<256>(v, lds);
  double sc = block_sum<256>(c, lds);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = sv;
    partial_n[blockIdx.x] = sc;
  }
}

// b[fluid] -= mean (doubly periodic compatibility, fb/core/_scipy_adapter.py:109-111), or the
// same on the solution (:153-154).
__global__ void __launch_bounds__(256) sub_mean_kernel(size_t N, const uint8_t* __restrict__ fl, double* __restrict__ a,
                                                       double mean) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N && !(fl[i] & PF_SOLID)) a[i] -= mean;
}

__global__ void __launch_bounds__(256) sub_mean_dev_kernel(size_t N, const uint8_t* __restrict__ fl,
                                                           double* __restrict__ a, const double* __restrict__ scal) {
  const double cnt = scal[S_BCNT];
  if (!(cnt > 0.0)) return;
  const double mean = scal[S_BSUM] / cnt;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N && !(fl[i] & PF_SOLID)) a[i] -= mean;
}

// Deterministic sum of n block partials into scal[slot].
__global__ void __launch_bounds__(256) sum_slot_kernel(const double* __restrict__ partial, int n, double* scal,
                                                       int slot, int ps) {
  __shared__ double lds[4];
  partial += (size_t)blockIdx.z * ps;
  scal += (size_t)blockIdx.z * kScalStride;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
  s = block_sum<256>(s, lds);
  if (threadIdx.x == 0) scal[slot] = s;
}

// Non-finite or out-of-range right-hand side: stop the CG before its first iteration.
__global__ void rhs_check_kernel(double* scal, double safe) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  scal += (size_t)blockIdx.z * kScalStride;
  const unsigned long long* st = reinterpret_cast<const unsigned long long*>(scal + S_COUNT);
  if (st[0] != 0ull) {
    scal[S_FAIL] = 1.0;
    scal[S_DONE] = 1.0;
  } else if (__longlong_as_double((long long)st[1]) > safe) {
    scal[S_FAIL] = 2.0;
    scal[S_DONE] = 1.0;
  }
}

__global__ void __launch_bounds__(256) to_f32_scaled(size_t N, const double* __restrict__ a, float* __restrict__ b,
                                                     double s) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N) b[i] = (float)(a[i] * s);
}

__global__ void __launch_bounds__(256) add_f32_scaled(size_t N, double* __restrict__ x, const float* __restrict__ e,
                                                      double s) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N) x[i] += (double)e[i] * s;
}

// Warm start: a grid already done before its first iteration (||b|| == 0, or a rejected
// right-hand side) returns x = 0, as SciPy returns b for b == 0.
__global__ void __launch_bounds__(256) zero_if_done_kernel(size_t N, double* __restrict__ x, const double* scal) {
  if (!CG_DONE(scal + (size_t)blockIdx.z * kScalStride)) return;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N) x[(size_t)blockIdx.z * N + i] = 0.0;
}

// Batched-solve staging in one launch per direction: grid z copies its N values between its own
// buffer (a pointer table in dpar: the right-hand sides first, then the solutions) and its slice
// of the stacked arrays. gather: rhs -> bflat and, with `with_x`, x -> xflat; else xflat -> x.
__global__ void __launch_bounds__(256) batch_stage_kernel(size_t N, int G, const double* __restrict__ tbl,
                                                          double* __restrict__ bflat, double* __restrict__ xflat,
                                                          int gather, int with_x) {
  const int g = blockIdx.z;
  const double* rhs = reinterpret_cast<const double* const*>(tbl)[g];
  double* x = reinterpret_cast<double* const*>(tbl)[G + g];
  const size_t o = (size_t)g * N;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t)gridDim.x * 256) {
    if (gather) {
      bflat[o + i] = rhs[i];
      if (with_x) xflat[o + i] = x[i];
    } else {
      x[i] = xflat[o + i];
    }
  }
}

__global__ void __launch_bounds__(256) fill_zero(size_t N, double* __restrict__ a) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N) a[i] = 0.0;
}

// Per-grid scaled copies of the refinement loop: grid blockIdx.z uses scale s[blockIdx.z].
__global__ void __launch_bounds__(256) to_f32_scaled_g(size_t N, const double* __restrict__ a, float* __restrict__ b,
                                                       const double* __restrict__ s) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t o = (size_t)blockIdx.z * N + i;
  if (i < N) b[o] = (float)(a[o] * s[blockIdx.z]);
}
__global__ void __launch_bounds__(256) add_f32_scaled_g(size_t N, double* __restrict__ x, const float* __restrict__ e,
                                                        const double* __restrict__ s) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t o = (size_t)blockIdx.z * N + i;
  if (i < N) x[o] += (double)e[o] * s[blockIdx.z];
}

// CG scalars of grid g = blockIdx.z: atol <- rtol[g] (scaled by ||b|| later), maxiter <-
// maxit[g] (par = [maxit x nb, rtol x nb]); a grid with maxit 0 ends before its first iteration.
__global__ void cg_init_kernel(double* scal, const double* __restrict__ par, int nb) {
  double* sc = scal + (size_t)blockIdx.z * kScalStride;
  if (threadIdx.x == S_COUNT) *reinterpret_cast<unsigned*>(sc + S_TICKET) = 0u;
  if (threadIdx.x >= S_COUNT) return;
  const int k = threadIdx.x;
  sc[k] = k == S_ATOL ? par[nb + blockIdx.z] : (k == S_MAXIT ? par[blockIdx.z] : 0.0);
}

// ---------------------------------------------------------------- host driver

Poisson::~Poisson() {
  for (auto& g : graphs_)
    if (g.exec) (void)graph_exec_free(g.exec);
  for (void* ptr : {(void*)flags, (void*)diag, (void*)ex, (void*)partial, (void*)scal, (void*)b64, (void*)r64,
                    (void*)bstage, (void*)xstage, (void*)dpar})
    if (ptr) (void)dev_free(ptr);
  if (host_scal) (void)host_free(host_scal);
  for (auto& e : par_ev)
    if (e) (void)hipEventDestroy(e);
  if (host_par) (void)host_free(host_par);
  auto freev = [](auto& v) {
    for (void* ptr : {(void*)v.r, (void*)v.z, (void*)v.p, (void*)v.q, (void*)v.x, (void*)v.b, (void*)v.work,
                      (void*)v.scratch, (void*)v.tri, (void*)v.tri_itf, (void*)v.tri_yF, (void*)v.tri_yE})
      if (ptr) (void)dev_free(ptr);
    free_plan(v.plan);
  };
  freev(v64);
  freev(v32);
}

template <class T>
int Poisson::alloc(CgVecs<T>& v) {
  if (v.r) return FB_OK;
  const size_t n = (size_t)cap * N;
  for (T** a : {&v.r, &v.z, &v.q, &v.x, &v.b, &v.work, &v.scratch}) FB_HIP(dev_alloc((void**)a, n * sizeof(T)));
  FB_HIP(dev_alloc((void**)&v.p, 2 * n * sizeof(T)));
  if (!periodic_x && !periodic_y) {
    int rc = make_plan<T>(v.plan, nx, stream);
    if (rc) return rc;
    return make_tri<T>(v);
  }
  return FB_OK;
}

// Factors, spikes and interface coefficients of the y-direction partition solver (long double
// on the host), in the [chunk type][quantity][row][mode] and [quantity][chunk][mode] layouts.
template <class T>
int Poisson::make_tri(CgVecs<T>& v) {
  const int L = kTriL;
  const int P = (ny + L - 1) / L;
  const int Ll = ny - (P - 1) * L;
  v.tri_Llast = Ll;
  v.tri_P = P;
  std::vector<T> fac((size_t)8 * L * nx, (T)0);
  std::vector<T> itf((size_t)6 * P * nx);
  const long double iy2 = 1.0L / ((long double)dy * (long double)dy);
  const long double a = -iy2;
  std::vector<double> exh(nx);
  for (int i = 0; i < nx; ++i)
    exh[i] = 2.0 * (1.0 - std::cos(M_PI * (double)(i + 1) / (double)(nx + 1))) * (1.0 / (dx * dx));
  std::vector<long double> m, al, be;
  long double ends[2][4];  // per chunk type: alpha first, beta first, alpha last, beta last
  for (int i = 0; i < nx; ++i) {
    const long double b = 2.0L * iy2 + (long double)exh[i];
    for (int part = 0; part < 2; ++part) {
      const int Lc = part == 0 ? L : Ll;
      auto at = [&](int q, int j) -> T& { return fac[((size_t)(part * 4 + q) * L + j) * nx + i]; };
      m.assign(Lc, 0.0L);
      al.assign(Lc, 0.0L);
      be.assign(Lc, 0.0L);
      m[0] = b;
      for (int j = 1; j < Lc; ++j) m[j] = b - a * a / m[j - 1];
      for (int j = 0; j < Lc; ++j) {
        at(0, j) = (T)(1.0L / m[j]);
        at(1, j) = (T)(a / m[j]);
      }
      // alpha: coupling to the value below the chunk (right-hand side -a in the first row)
      for (int j = 0; j < Lc; ++j) al[j] = ((j == 0 ? -a : 0.0L) - (j ? a * al[j - 1] : 0.0L)) / m[j];
      for (int j = Lc - 2; j >= 0; --j) al[j] = al[j] - (a / m[j]) * al[j + 1];
      // beta: coupling to the value above the chunk (right-hand side -a in the last row)
      for (int j = 0; j < Lc; ++j) be[j] = ((j == Lc - 1 ? -a : 0.0L) - (j ? a * be[j - 1] : 0.0L)) / m[j];
      for (int j = Lc - 2; j >= 0; --j) be[j] = be[j] - (a / m[j]) * be[j + 1];
      for (int j = 0; j < Lc; ++j) {
        at(2, j) = (T)al[j];
        at(3, j) = (T)be[j];
      }
      ends[part][0] = al[0];
      ends[part][1] = be[0];
      ends[part][2] = al[Lc - 1];
      ends[part][3] = be[Lc - 1];
    }
    // interface recurrence coefficients (they do not depend on the right-hand side)
    long double qq = 0.0L;
    const size_t PQ = (size_t)P * nx;
    for (int k = 0; k < P; ++k) {
      const long double* e = ends[k == P - 1 ? 1 : 0];
      const long double aF = e[0], bF = e[1], aE = e[2], bE = e[3];
      const long double den = 1.0L - aF * qq;
      const long double sv = bF / den;
      const long double mm = aE * qq;
      const size_t o = (size_t)k * nx + i;
      itf[o] = (T)(1.0L / den);
      itf[PQ + o] = (T)(aF / den);
      itf[2 * PQ + o] = (T)(aE + mm * aF / den);
      itf[3 * PQ + o] = (T)(mm / den);
      itf[4 * PQ + o] = (T)sv;
      qq = mm * sv + bE;
      itf[5 * PQ + o] = (T)qq;
    }
  }
  FB_HIP(dev_alloc((void**)&v.tri, fac.size() * sizeof(T)));
  FB_HIP(dev_alloc((void**)&v.tri_itf, itf.size() * sizeof(T)));
  FB_HIP(dev_alloc((void**)&v.tri_yF, (size_t)cap * P * nx * sizeof(T)));
  FB_HIP(dev_alloc((void**)&v.tri_yE, (size_t)cap * P * nx * sizeof(T)));
  FB_HIP(hipMemcpyAsync(v.tri, fac.data(), fac.size() * sizeof(T), hipMemcpyHostToDevice, stream));
  FB_HIP(hipMemcpyAsync(v.tri_itf, itf.data(), itf.size() * sizeof(T), hipMemcpyHostToDevice, stream));
  FB_HIP(hipStreamSynchronize(stream));
  return FB_OK;
}

int Poisson::setup(hipStream_t s, int nx_, int ny_, double dx_, double dy_, int px, int py, const uint8_t* d_solid,
                   int capacity) {
  stream = s;
  if (flags == nullptr || nx_ != nx || ny_ != ny || px != periodic_x || py != periodic_y) {
    if (flags) {
      set_error("Poisson solver cannot change shape after set-up");
      return FB_ERR_STATE;
    }
    FB_REQUIRE(capacity >= 1 && (capacity == 1 || (!px && !py)), FB_ERR_ARG,
               "batched Poisson solves need a non-periodic box");
    cap = capacity;
    nx = nx_;
    ny = ny_;
    dx = dx_;
    dy = dy_;
    periodic_x = px;
    periodic_y = py;
    N = (size_t)nx * ny;
    nblocks = (int)((N + 255) / 256);
    ps = 2 * (nblocks + ny + 64);
    FB_HIP(dev_alloc((void**)&flags, (size_t)cap * N));
    FB_HIP(dev_alloc((void**)&diag, (size_t)cap * N * sizeof(double)));
    FB_HIP(dev_alloc((void**)&partial, (size_t)cap * ps * sizeof(double)));
    FB_HIP(dev_alloc((void**)&scal, (size_t)cap * kScalStride * sizeof(double)));
    FB_HIP(host_alloc((void**)&host_scal, (size_t)cap * kScalStride * sizeof(double)));
    FB_HIP(dev_alloc((void**)&dpar, (size_t)2 * cap * sizeof(double)));
    FB_HIP(host_alloc((void**)&host_par, (size_t)kParSlots * 2 * cap * sizeof(double)));
    FB_HIP(dev_alloc((void**)&b64, (size_t)cap * N * sizeof(double)));
    FB_HIP(dev_alloc((void**)&r64, (size_t)cap * N * sizeof(double)));
    if (cap > 1) {
      FB_HIP(dev_alloc((void**)&bstage, (size_t)cap * N * sizeof(double)));
      FB_HIP(dev_alloc((void**)&xstage, (size_t)cap * N * sizeof(double)));
    }
    FB_HIP(dev_alloc((void**)&ex, nx * sizeof(double)));
    std::vector<double> exh(nx);
    for (int i = 0; i < nx; ++i)
      exh[i] = 2.0 * (1.0 - std::cos(M_PI * (double)(i + 1) / (double)(nx + 1))) * (1.0 / (dx * dx));
    FB_HIP(hipMemcpyAsync(ex, exh.data(), nx * sizeof(double), hipMemcpyHostToDevice, stream));
    FB_HIP(hipStreamSynchronize(stream));
    int rc = alloc(v64);
    if (rc) return rc;
  }
  if (d_solid) {
    hipLaunchKernelGGL(setup_kernel, dim3(nblocks), dim3(256), 0, stream, nx, ny, d_solid, px, py, 1.0 / (dx * dx),
                       1.0 / (dy * dy), flags, diag);
    FB_HIP(hipGetLastError());
    static std::atomic<unsigned long long> tags{0};
    version = ++tags;
  }
  return FB_OK;
}

int Poisson::load_slot(int k, const Poisson& src) {
  FB_REQUIRE(k >= 0 && k < cap && src.nx == nx && src.ny == ny && src.flags, FB_ERR_ARG,
             "slot %d does not fit a %dx%d grid", k, src.nx, src.ny);
  FB_HIP(hipMemcpyAsync(flags + (size_t)k * N, src.flags, N, hipMemcpyDeviceToDevice, stream));
  FB_HIP(hipMemcpyAsync(diag + (size_t)k * N, src.diag, N * sizeof(double), hipMemcpyDeviceToDevice, stream));
  return FB_OK;
}

int Poisson::reduce(const double* pin, int n, int mode) {
  Launcher L{stream};
  return reduce(L, pin, n, mode);
}
int Poisson::reduce(Launcher& L, const double* pin, int n, int mode) {
  L(cg_finalize_kernel, dim3(1, 1, nb_), dim3(256), 0, pin, n, scal, mode, ps);
  FB_HIP(L.status());
  return FB_OK;
}

int Poisson::read_scalars() {
  FB_HIP(hipMemcpyAsync(host_scal, scal, ((size_t)(nb_ - 1) * kScalStride + S_COUNT) * sizeof(double),
                        hipMemcpyDeviceToHost, stream));
  FB_HIP(hipStreamSynchronize(stream));
  return FB_OK;
}

// Per-grid launch parameters to dpar. The copy reads pinned memory when the stream reaches it,
// so successive calls use successive host slots of a ring, and a slot is rewritten only once the
// copy that last read it has run (its event).
int Poisson::put_par(const double* v, int count) {
  const unsigned k = par_slot_++ % kParSlots;
  if (par_ev[k]) FB_HIP(hipEventSynchronize(par_ev[k]));
  else FB_HIP(hipEventCreateWithFlags(&par_ev[k], hipEventDisableTiming));
  double* h = host_par + (size_t)k * 2 * cap;
  for (int g = 0; g < count; ++g) h[g] = v[g];
  FB_HIP(hipMemcpyAsync(dpar, h, count * sizeof(double), hipMemcpyHostToDevice, stream));
  FB_HIP(hipEventRecord(par_ev[k], stream));
  return FB_OK;
}

template <class T>
int Poisson::precond(CgVecs<T>& v, const T* rin, T* zout) {
  Launcher L{stream};
  return precond(L, v, rin, zout);
}

template <class T>
int Poisson::precond(Launcher& L, CgVecs<T>& v, const T* rin, T* zout) {
  if (periodic_x || periodic_y) {
    L(jacobi_kernel<T>, dim3(nblocks), dim3(256), 0, N, diag, rin, zout, partial, scal,
      (double)std::numeric_limits<double>::epsilon());
    FB_HIP(L.status());
    return reduce(L, partial, nblocks, 1);
  }
  if constexpr (sizeof(T) == 8) {
    if (precond32_) return dst_precond<float, double>(L, v32, rin, zout);
  }
  return dst_precond<T, T>(L, v, rin, zout);
}

template <class T, class TIO>
int Poisson::dst_precond(Launcher& L, CgVecs<T>& v, const TIO* rin, TIO* zout) {
  const DstPlan& P = *v.plan;
  const int tp = sizeof(T) == 8 ? FB_F64 : FB_F32;
  // points per thread (16, or 8 for twice the threads per row) and rows per block, as tuned for
  // this shape and batch size (or any batch size). By default 8 points per thread for transform
  // lengths 512 to 4096 (measured faster there, the latency-bound passes get twice the waves),
  // and several short rows per block, but never fewer blocks than about two per CU.
  auto tuned = [&](int param) {
    const int t = tune_get(param, nx, ny, tp, nb_);
    return t >= 1 ? t : tune_get(param, nx, ny, tp, 0);
  };
  int ppt = tuned(TUNE_DST_PPT);
  if (ppt < 1) ppt = (P.M >= 512 && P.M <= 4096) ? 8 : 16;
  if (ppt != 8 || P.M / 8 > 512) ppt = 16;
  const int nthr = P.M / ppt;
  int rows = tuned(TUNE_DST_ROWS);
  if (rows < 1 || rows * nthr > 512 || (size_t)rows * lds_row(P.M) * sizeof(cplx<T>) > 96 * 1024)
    rows = std::max(1, std::min(256 / nthr, ny * nb_ / 512));
  const dim3 dg((unsigned)((ny + rows - 1) / rows), 1, nb_);
  const size_t lds = (size_t)rows * lds_row(P.M) * sizeof(cplx<T>) + (size_t)rows * (P.M / 16) * sizeof(double);
  // the FP64 transform gains from the fused input and output passes; the FP32 one from the fused
  // output only (a fused input takes it from 59 to 94 registers, and was measured slower)
  constexpr int fuse = sizeof(T) == 8 ? 3 : 2;
  auto* kern_in = ppt == 8 ? dst_rows_kernel<T, TIO, T, 8, fuse, 0> : dst_rows_kernel<T, TIO, T, 16, fuse, 0>;
  auto* kern_out = ppt == 8 ? dst_rows_kernel<T, T, TIO, 8, fuse, 1> : dst_rows_kernel<T, T, TIO, 16, fuse, 1>;
  L(kern_in, dg, dim3(rows * nthr), lds, nx, ny, rin, flags, (const T*)v.work, v.work, (const cplx<T>*)P.chirp,
    (const cplx<T>*)P.fb, (const cplx<T>*)P.tw, P.M, P.logM, rows, (T)P.scale, partial, scal, ps);
  FB_HIP(L.status());
  const T a = (T)(-(1.0 / (dy * dy)));
  if (v.tri_P <= kTriFusedMaxP) {
    L(tri_solve_kernel<T>, dim3((nx + 15) / 16, 1, nb_), dim3(16, tri_threads(v.tri_P)), 0, nx, ny, v.work,
      (const T*)v.tri, v.tri_Llast, v.tri_P, a, (const T*)v.tri_itf, v.tri_yF, v.tri_yE, scal);
  } else {
    const dim3 tb(64, 4), tg((nx + 63) / 64, (v.tri_P + 3) / 4, nb_);
    L(tri_chunk_kernel<T>, tg, tb, 0, nx, ny, v.work, (const T*)v.tri, v.tri_Llast, v.tri_P, a, v.tri_yF, v.tri_yE,
      scal);
    // segments of up to 8 chunks run from registers (columns of up to 128 chunks)
    auto* itf_kern = (v.tri_P + 15) / 16 <= 8 ? tri_interface_kernel<T, 8> : tri_interface_kernel<T, 0>;
    L(itf_kern, dim3((nx + 15) / 16, 1, nb_), dim3(16, 16), 0, nx, v.tri_P, (const T*)v.tri_itf, v.tri_yF, v.tri_yE,
      scal);
    L(tri_spike_kernel<T>, tg, tb, 0, nx, ny, v.work, (const T*)v.tri, v.tri_Llast, v.tri_P, (const T*)v.tri_yF,
      (const T*)v.tri_yE, scal);
  }
  FB_HIP(L.status());
  // r.z is folded by a launch of its own: folding it in the transform's last block was measured
  // slower (256 x 128 FP64: 12.8 us against 7.3 + 3.9), the release fence of every block writing
  // back its XCD's L2 lines of z
  L(kern_out, dg, dim3(rows * nthr), lds, nx, ny, (const T*)v.work, flags, rin, zout, (const cplx<T>*)P.chirp,
    (const cplx<T>*)P.fb, (const cplx<T>*)P.tw, P.M, P.logM, rows, (T)P.scale, partial, scal, ps);
  FB_HIP(L.status());
  return reduce(L, partial, ny, 1);
}

template <class T>
int Poisson::pcg(CgVecs<T>& v, const T* b, T* x, const double* rtol, const int* maxit, bool check_rhs,
                 int first_chunk, int* iterations, int* info, int* failure, bool warm) {
  const size_t n = (size_t)nb_ * N;
  if constexpr (sizeof(T) != 8) warm = false;
  if (!warm) {
    FB_HIP(hipMemsetAsync(x, 0, n * sizeof(T), stream));
    FB_HIP(hipMemcpyAsync(v.r, b, n * sizeof(T), hipMemcpyDeviceToDevice, stream));
  }
  int maxall = 0;
  {
    double par[2 * kMaxBatch];
    for (int g = 0; g < nb_; ++g) {
      par[g] = (double)maxit[g];
      par[nb_ + g] = rtol[g];
      maxall = std::max(maxall, maxit[g]);
    }
    int rc = put_par(par, 2 * nb_);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(cg_init_kernel, dim3(1, 1, nb_), dim3(64), 0, stream, scal, (const double*)dpar, nb_);
  // ||b||^2 (r = b when starting from zero) sets the absolute tolerance
  hipLaunchKernelGGL(dot_partial_kernel<T>, dim3(nblocks, 1, nb_), dim3(256), 0, stream, warm ? b : (const T*)v.r,
                     warm ? b : (const T*)v.r, N, partial, ps);
  FB_HIP(hipGetLastError());
  int rc = reduce(partial, nblocks, 4);
  if (rc) return rc;
  if (check_rhs) {
    // rejections of fb/core/_scipy_adapter.py:113-120, decided on the device
    const double safe = 0.25 * std::sqrt(std::numeric_limits<double>::max() / (double)std::max<size_t>(N, 1));
    hipLaunchKernelGGL(rhs_check_kernel, dim3(1, 1, nb_), dim3(64), 0, stream, scal, safe);
    FB_HIP(hipGetLastError());
  }
  if constexpr (sizeof(T) == 8) {
    if (warm) {
      // r0 = b - A x0 and the residual tests on it; grids done already return x = 0
      hipLaunchKernelGGL(zero_if_done_kernel, dim3(nblocks, 1, nb_), dim3(256), 0, stream, N, x,
                         (const double*)scal);
      hipLaunchKernelGGL(residual_kernel, dim3(nblocks, 1, nb_), dim3(256), 0, stream, nx, ny, flags, diag, b, x,
                         v.r, partial, 1.0 / (dx * dx), 1.0 / (dy * dy), ps);
      FB_HIP(hipGetLastError());
      if ((rc = reduce(partial, nblocks, 5))) return rc;
    }
  }
  int launched = 0;
  // the first poll comes after about as many iterations as the previous solve needed
  int chunk = std::min(std::max(first_chunk, 4), 64);
  for (;;) {
    const int nn = std::min(chunk, maxall - launched);
    if (nn > 0 && (rc = iterations_chunk(v, x, nn, launched & 1))) return rc;
    launched += std::max(nn, 0);
    if ((rc = read_scalars())) return rc;
    bool all = true;
    for (int g = 0; g < nb_; ++g) all = all && host_scal[(size_t)g * kScalStride + S_DONE] != 0.0;
    if (all || launched >= maxall) break;
    chunk = std::min(2 * chunk, 32);
  }
  for (int g = 0; g < nb_; ++g) {
    const double* h = host_scal + (size_t)g * kScalStride;
    iterations[g] = (int)h[S_ITER];
    info[g] = (h[S_DONE] != 0.0) ? (int)h[S_INFO] : maxit[g];
    failure[g] = (int)h[S_FAIL];
  }
  return FB_OK;
}

// A chunk of n CG iterations (with the preconditioner). Between host polls the
// iterations only read and write device memory, so a chunk is built once as a HIP graph
// per (precision, solution buffer, length, parity, grids) and replayed; every kernel returns at
// once when its grid's device-side done flag is set.
template <class T>
int Poisson::iterations_chunk(CgVecs<T>& v, T* x, int n, int parity) {
  if (n <= 0) return FB_OK;
  const GraphKey key{sizeof(T), (const void*)x, n, parity, nb_, precond32_ ? 1 : 0};
  for (auto& g : graphs_)
    if (g.key == key) {
      FB_HIP(hipGraphLaunch(g.exec, stream));
      return FB_OK;
    }
  Launcher direct{stream};
  if (!use_graphs_) return enqueue_iterations(direct, v, x, n, parity);
  // a graph is built for a chunk shape the second time it comes up: a one-off solve (such as a
  // set-up projection) does not pay for building and instantiation
  bool seen = false;
  for (auto& k : seen_)
    if (k == key) seen = true;
  if (!seen) {
    seen_.push_back(key);
    return enqueue_iterations(direct, v, x, n, parity);
  }
  hipGraph_t graph = nullptr;
  FB_HIP(hipGraphCreate(&graph, 0));
  Launcher L{stream, graph};
  int rc = enqueue_iterations(L, v, x, n, parity);
  if (rc) {
    (void)hipGraphDestroy(graph);
    return rc;
  }
  hipGraphExec_t exec = nullptr;
  hipError_t e;
  {
    std::lock_guard<std::recursive_mutex> guard(hip_lock());
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  }
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    set_error("CG graph instantiation failed: %s", hipGetErrorString(e));
    return FB_ERR_HIP;
  }
  graphs_.push_back({key, exec});
  FB_HIP(hipGraphLaunch(exec, stream));
  return FB_OK;
}

template <class T>
int Poisson::enqueue_iterations(Launcher& L, CgVecs<T>& v, T* x, int n, int parity) {
  const T ix2 = (T)(1.0 / (dx * dx)), iy2 = (T)(1.0 / (dy * dy));
  const size_t pn = (size_t)cap * N;
  T* pa = parity ? v.p + pn : v.p;
  T* pb = parity ? v.p : v.p + pn;
  const dim3 grid(nblocks, 1, nb_);
  const dim3 pqgrid((unsigned)((nx + kPqW - 1) / kPqW), (unsigned)((ny + pq_rows(nx, ny) - 1) / pq_rows(nx, ny)), nb_);
  FB_REQUIRE(pq_tiles(nx, ny) <= ps, FB_ERR_ARG, "operator tiles exceed the partial-sum slots");
  int rc;
  for (int c = 0; c < n; ++c) {
    if ((rc = precond(L, v, v.r, v.z))) return rc;
    // on small grids each of these folds its partials in its last block (pq and alpha; ||r||^2
    // and the check); otherwise a fold launch follows each
    const int fold = nblocks <= kFoldBlocks ? 1 : 0;
    const int ntiles = pq_tiles(nx, ny);
    const int pfold = ntiles <= kFoldBlocks ? 1 : 0;
    L(pq_rows(nx, ny) == 16 ? cg_pq_kernel<T, 16> : cg_pq_kernel<T, 4>, pqgrid, dim3(256), 0, nx, ny, flags, v.z, pa, pb,
      v.q, partial, scal, ix2, iy2, 1.0 / (dx * dx), 1.0 / (dy * dy), periodic_x ? 1 : 0, periodic_y ? 1 : 0, ps,
      pfold);
    FB_HIP(L.status());
    if (!pfold && (rc = reduce(L, partial, ntiles, 2))) return rc;
    L(cg_update_kernel<T>, grid, dim3(256), 0, N, x, v.r, pb, v.q, partial, scal, ps, fold);
    FB_HIP(L.status());
    if (!fold && (rc = reduce(L, partial, nblocks, 3))) return rc;
    std::swap(pa, pb);
  }
  return FB_OK;
}

int Poisson::sum_to(const double* pin, int n, int slot) {
  hipLaunchKernelGGL(sum_slot_kernel, dim3(1, 1, nb_), dim3(256), 0, stream, pin, n, scal, slot, ps);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

int Poisson::solve(const double* d_rhs, double* d_x, double tolerance, int maxit, int mode, PoissonReport* rep) {
  nb_ = 1;
  return run(d_rhs, d_x, tolerance, maxit, mode, rep);
}

int Poisson::solve_batch(int nb, const double* const* rhs, double* const* x, double tolerance, int maxit, int mode,
                         PoissonReport* reps) {
  FB_REQUIRE(nb >= 1 && nb <= cap && nb <= kMaxBatch, FB_ERR_ARG, "batch of %d grids exceeds the capacity %d", nb,
             cap);
  if (nb == 1 && cap == 1) {
    int rc = solve(rhs[0], x[0], tolerance, maxit, mode, reps);
    if (rc) return rc;
    FB_HIP(hipStreamSynchronize(stream));
    finish(reps);
    return FB_OK;
  }
  nb_ = nb;
  // one staging launch each way for the whole batch (its pointer table staged like the launch
  // parameters): the grids' right-hand sides (and warm-start solutions) in, the solutions out
  double tbl[2 * kMaxBatch];
  for (int g = 0; g < nb; ++g) {
    memcpy(&tbl[g], &rhs[g], sizeof(double));
    memcpy(&tbl[nb + g], &x[g], sizeof(double));
  }
  const dim3 sg((unsigned)std::min<size_t>((N + 255) / 256, 2048), 1, (unsigned)nb);
  int rc = put_par(tbl, 2 * nb);
  if (rc) return rc;
  hipLaunchKernelGGL(batch_stage_kernel, sg, dim3(256), 0, stream, N, nb, (const double*)dpar, bstage, xstage, 1,
                     (mode & kWarmStart) ? 1 : 0);
  FB_HIP(hipGetLastError());
  rc = run(bstage, xstage, tolerance, maxit, mode, reps);
  if (rc) return rc;
  if ((rc = put_par(tbl, 2 * nb))) return rc;
  hipLaunchKernelGGL(batch_stage_kernel, sg, dim3(256), 0, stream, N, nb, (const double*)dpar, bstage, xstage, 0, 0);
  FB_HIP(hipGetLastError());
  FB_HIP(hipStreamSynchronize(stream));
  for (int g = 0; g < nb; ++g) finish(reps + g, g);
  return FB_OK;
}

// The solve of nb_ grids whose right-hand sides sit N apart at d_rhs; solutions go N apart to d_x.
int Poisson::run(const double* d_rhs, double* d_x, double tolerance, int maxit, int mode, PoissonReport* reps) {
  const int G = nb_;
  for (int g = 0; g < G; ++g) reps[g] = PoissonReport{};
  const int nb = nblocks;
  const size_t n = (size_t)G * N;
  const unsigned nbz = (unsigned)((n + 255) / 256);
  FB_HIP(hipMemsetAsync(scal, 0, (size_t)G * kScalStride * sizeof(double), stream));
  hipLaunchKernelGGL(rhs_prep_kernel, dim3(nb, 1, G), dim3(256), 0, stream, N, d_rhs, flags, b64, partial,
                     partial + nb, reinterpret_cast<unsigned long long*>(scal + S_COUNT), ps);
  FB_HIP(hipGetLastError());
  const bool singular = periodic_x && periodic_y;
  int rc;
  if (singular) {
    // compatible right-hand side: remove the fluid mean (fb/core/_scipy_adapter.py:108-112)
    if ((rc = sum_to(partial, nb, S_BSUM)) || (rc = sum_to(partial + nb, nb, S_BCNT))) return rc;
    hipLaunchKernelGGL(sub_mean_dev_kernel, dim3(nb), dim3(256), 0, stream, N, flags, b64, (const double*)scal);
    FB_HIP(hipGetLastError());
  }
  const double rtol = 0.9 * tolerance;
  const double ix2 = 1.0 / (dx * dx), iy2 = 1.0 / (dy * dy);
  const bool warm = (mode & kWarmStart) != 0;
  mode &= kPrecisionMask;
  FB_REQUIRE(!(warm && mode == kRefine), FB_ERR_ARG, "a warm-started pressure solve needs FP64 CG (mode 0 or 2)");
  pending_failure_ = 0;
  int its[kMaxBatch], info[kMaxBatch], fail[kMaxBatch], mi[kMaxBatch];
  double rt[kMaxBatch];
  if (mode != kRefine) {
    for (int g = 0; g < G; ++g) {
      mi[g] = maxit;
      rt[g] = rtol;
    }
    // kPrecond32: the same FP64 PCG, its preconditioner applied with the FP32 tables
    if (mode == kPrecond32 && (rc = alloc(v32))) return rc;
    precond32_ = mode == kPrecond32;
    rc = pcg<double>(v64, b64, d_x, rt, mi, true, last_iterations_ + 1, its, info, fail, warm);
    precond32_ = false;
    if (rc) return rc;
    int mx = 0;
    for (int g = 0; g < G; ++g) {
      reps[g].iterations = its[g];
      reps[g].info = info[g];
      reps[g].failure = fail[g];
      mx = std::max(mx, its[g]);
    }
    last_iterations_ = mx;
  } else {
    // FP32 PCG sweeps on the scaled FP64 residual, FP64 correction and residual.
    if ((rc = alloc(v32))) return rc;
    FB_HIP(hipMemcpyAsync(host_scal, scal, (size_t)G * kScalStride * sizeof(double), hipMemcpyDeviceToHost,
                          stream));
    FB_HIP(hipStreamSynchronize(stream));
    const double safe = 0.25 * std::sqrt(std::numeric_limits<double>::max() / (double)std::max<size_t>(N, 1));
    bool failed[kMaxBatch];
    for (int g = 0; g < G; ++g) {
      const unsigned long long* st = reinterpret_cast<const unsigned long long*>(host_scal + (size_t)g * kScalStride +
                                                                                 S_COUNT);
      double bmax;
      {
        long long bits = (long long)st[1];
        memcpy(&bmax, &bits, sizeof(bmax));
      }
      failed[g] = st[0] != 0 || bmax > safe;
      if (failed[g]) reps[g].failure = st[0] != 0 ? 1 : 2;
    }
    hipLaunchKernelGGL(fill_zero, dim3(nbz), dim3(256), 0, stream, n, d_x);
    FB_HIP(hipGetLastError());
    FB_HIP(hipMemcpyAsync(r64, b64, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(dot_partial_kernel<double>, dim3(nb, 1, G), dim3(256), 0, stream, b64, b64, N, partial, ps);
    if ((rc = reduce(partial, nb, 0))) return rc;
    if ((rc = read_scalars())) return rc;
    double bn[kMaxBatch], rn[kMaxBatch];
    int total[kMaxBatch];
    for (int g = 0; g < G; ++g) {
      bn[g] = std::sqrt(host_scal[(size_t)g * kScalStride + S_DOT]);
      rn[g] = bn[g];
      total[g] = 0;
      reps[g].info = maxit;
    }
    for (int outer = 0; outer < 12; ++outer) {
      bool act[kMaxBatch], any = false;
      for (int g = 0; g < G; ++g) {
        act[g] = false;
        if (failed[g]) continue;
        if (bn[g] == 0.0 || rn[g] < rtol * bn[g]) {
          reps[g].info = 0;
          continue;
        }
        act[g] = total[g] < maxit;
        any = any || act[g];
      }
      if (!any) break;
      double sc[kMaxBatch];
      for (int g = 0; g < G; ++g) {
        sc[g] = act[g] ? 1.0 / rn[g] : 0.0;
        mi[g] = act[g] ? maxit - total[g] : 0;
        // inner tolerance: the reduction still missing (with a factor 2 of margin), but no tighter
        // than a relative reduction FP32 reaches reliably
        rt[g] = act[g] ? std::min(0.5, std::max(2.0e-5, 0.5 * rtol * bn[g] / rn[g])) : 1.0;
      }
      if ((rc = put_par(sc, G))) return rc;
      hipLaunchKernelGGL(to_f32_scaled_g, dim3(nb, 1, G), dim3(256), 0, stream, N, r64, v32.b, (const double*)dpar);
      FB_HIP(hipGetLastError());
      if ((rc = pcg<float>(v32, v32.b, v32.x, rt, mi, false, last_inner_ + 1, its, info, fail))) return rc;
      int mx = 0;
      for (int g = 0; g < G; ++g) {
        if (!act[g]) {
          sc[g] = 0.0;
          continue;
        }
        mx = std::max(mx, its[g]);
        total[g] += its[g];
        reps[g].refinements = outer + 1;
        if (fail[g]) {
          reps[g].failure = fail[g];
          failed[g] = true;
          sc[g] = 0.0;
        } else {
          sc[g] = rn[g];
        }
      }
      last_inner_ = mx;
      if ((rc = put_par(sc, G))) return rc;
      hipLaunchKernelGGL(add_f32_scaled_g, dim3(nb, 1, G), dim3(256), 0, stream, N, d_x, v32.x, (const double*)dpar);
      FB_HIP(hipGetLastError());
      hipLaunchKernelGGL(residual_kernel, dim3(nb, 1, G), dim3(256), 0, stream, nx, ny, flags, diag, b64, d_x, r64,
                         partial, ix2, iy2, ps);
      FB_HIP(hipGetLastError());
      if ((rc = reduce(partial, nb, 0))) return rc;
      if ((rc = read_scalars())) return rc;
      for (int g = 0; g < G; ++g) {
        if (!act[g] || failed[g]) continue;
        rn[g] = std::sqrt(host_scal[(size_t)g * kScalStride + S_DOT]);
        if (!std::isfinite(rn[g])) {
          reps[g].failure = 4;
          failed[g] = true;
        }
      }
    }
    for (int g = 0; g < G; ++g) {
      if (reps[g].info != 0 && !failed[g] && rn[g] < rtol * bn[g]) reps[g].info = 0;
      reps[g].iterations = total[g];
    }
  }
  if (singular) {
    hipLaunchKernelGGL(rhs_prep_kernel, dim3(nb), dim3(256), 0, stream, N, d_x, flags, r64, partial, partial + nb,
                       reinterpret_cast<unsigned long long*>(scal + S_COUNT), ps);
    if ((rc = sum_to(partial, nb, S_BSUM)) || (rc = sum_to(partial + nb, nb, S_BCNT))) return rc;
    hipLaunchKernelGGL(sub_mean_dev_kernel, dim3(nb), dim3(256), 0, stream, N, flags, d_x, (const double*)scal);
    FB_HIP(hipGetLastError());
  }
  // true relative residual ||b - A x|| / max(||b||, eps) in FP64, read with the caller's next sync
  hipLaunchKernelGGL(residual_kernel, dim3(nb, 1, G), dim3(256), 0, stream, nx, ny, flags, diag, b64, d_x, r64,
                     partial, ix2, iy2, ps);
  FB_HIP(hipGetLastError());
  if ((rc = sum_to(partial, nb, S_RES2))) return rc;
  hipLaunchKernelGGL(dot_partial_kernel<double>, dim3(nb, 1, G), dim3(256), 0, stream, b64, b64, N, partial, ps);
  FB_HIP(hipGetLastError());
  if ((rc = sum_to(partial, nb, S_B2))) return rc;
  FB_HIP(hipMemcpyAsync(host_scal, scal, (size_t)G * kScalStride * sizeof(double), hipMemcpyDeviceToHost, stream));
  return FB_OK;
}

// Average duration of one preconditioner application to nb grids (timing for the autotuner; the
// vectors hold whatever they hold, which does not change the work).
int Poisson::time_precond(bool f32, int nb, int reps, double* ms) {
  FB_REQUIRE(nb >= 1 && nb <= cap && reps >= 1 && !periodic_x && !periodic_y, FB_ERR_ARG, "bad timing request");
  nb_ = nb;
  int rc = f32 ? alloc(v32) : alloc(v64);
  if (rc) return rc;
  FB_HIP(hipMemsetAsync(scal, 0, (size_t)nb * kScalStride * sizeof(double), stream));
  FB_HIP(hipMemsetAsync(flags, 0, (size_t)nb * N, stream));
  hipEvent_t e0, e1;
  FB_HIP(hipEventCreate(&e0));
  FB_HIP(hipEventCreate(&e1));
  auto once = [&]() { return f32 ? precond(v32, v32.r, v32.z) : precond(v64, v64.r, v64.z); };
  if ((rc = once())) return rc;  // warm-up
  FB_HIP(hipEventRecord(e0, stream));
  for (int k = 0; k < reps && rc == FB_OK; ++k) rc = once();
  FB_HIP(hipEventRecord(e1, stream));
  FB_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  FB_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / reps;
  return rc;
}

void Poisson::finish(PoissonReport* rep, int g) const {
  const double* h = host_scal + (size_t)g * kScalStride;
  const double rnorm = std::sqrt(h[S_RES2]);
  const double bnorm = std::sqrt(h[S_B2]);
  rep->relative_residual = rnorm / std::max(bnorm, std::numeric_limits<double>::epsilon());
}

}  // namespace fb
