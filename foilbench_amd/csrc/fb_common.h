// Shared runtime pieces of libfoilbench_amd: error state, HIP checks, device helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/foilbench_amd.h"

namespace fb {

// Thread-local last error message returned by fb_last_error().
void set_error(const char* fmt, ...);
void clear_error();

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  int xcd_count = 8;
};

}  // namespace fb

#define FB_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t fb_e_ = (call);                                                         \
    if (fb_e_ != hipSuccess) {                                                         \
      fb::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(fb_e_), __FILE__, \
                    __LINE__);                                                         \
      return FB_ERR_HIP;                                                               \
    }                                                                                  \
  } while (0)

#define FB_REQUIRE(cond, code, ...)  \
  do {                               \
    if (!(cond)) {                   \
      fb::set_error(__VA_ARGS__);    \
      return (code);                 \
    }                                \
  } while (0)

namespace fb {

// Order-preserving map of doubles onto unsigned 64-bit keys, for atomic min/max.
__device__ __forceinline__ unsigned long long dkey(double v) {
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double dkey_inv(unsigned long long k) {
  unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
  double v;
  memcpy(&v, &b, sizeof(v));
  return v;
}

// Max/min into a shared key with a read first: once the key holds the running extreme, most
// waves skip the atomic, which keeps same-address contention low on large grids.
__device__ __forceinline__ void key_max(unsigned long long* p, unsigned long long k) {
  if (k > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(p, k);
}
__device__ __forceinline__ void key_min(unsigned long long* p, unsigned long long k) {
  if (k < __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(p, k);
}

__device__ __forceinline__ int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// Wavefront (64-lane) reductions via DPP-backed shuffles.
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// XCD-aware block remap: blocks are dispatched round-robin over the XCDs; give each XCD
// a contiguous range of tiles so stencil neighbours share that XCD's L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nblocks) {
  const unsigned nx = 8;
  if (nblocks < nx * 4) return bid;
  unsigned per = nblocks / nx;
  unsigned full = per * nx;
  if (bid >= full) return bid;
  return (bid % nx) * per + bid / nx;
}

}  // namespace fb
