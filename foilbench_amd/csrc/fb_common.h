// Shared runtime pieces of libfoilbench_amd: error state, HIP checks, device helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>

#include "../../include/foilbench_amd.h"

namespace fb {

// Calls that may synchronise the device (allocation, free, pinned-host allocation, graph
// instantiation and destruction, stream creation) are serialised by this lock, so that handles
// can be created and freed from several threads while the request loop runs.
std::recursive_mutex& hip_lock();

// Device memory through a block cache: a freed block goes back to a per-size free list instead of
// hipFree (which synchronises the whole device and would stall every other stream), and the next
// allocation of that size class takes it. A block is freed only after the work of its owner has
// completed (the owners synchronise their streams first), so any stream may reuse it.
hipError_t cached_alloc(void** p, size_t bytes);
hipError_t cached_free(void* p);
// Returns the cached free blocks to HIP (for example before measuring free memory).
void release_cached_blocks();

template <class T>
inline hipError_t dev_alloc(T** p, size_t bytes) {
  return cached_alloc((void**)p, bytes);
}
inline hipError_t dev_free(void* p) { return cached_free(p); }
template <class T>
inline hipError_t host_alloc(T** p, size_t bytes) {
  std::lock_guard<std::recursive_mutex> g(hip_lock());
  return hipHostMalloc((void**)p, bytes, 0);
}
inline hipError_t host_free(void* p) {
  std::lock_guard<std::recursive_mutex> g(hip_lock());
  return hipHostFree(p);
}
inline hipError_t graph_exec_free(hipGraphExec_t e) {
  std::lock_guard<std::recursive_mutex> g(hip_lock());
  return hipGraphExecDestroy(e);
}
inline hipError_t stream_create(hipStream_t* s) {
  std::lock_guard<std::recursive_mutex> g(hip_lock());
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}
inline hipError_t stream_free(hipStream_t s) {
  std::lock_guard<std::recursive_mutex> g(hip_lock());
  return hipStreamDestroy(s);
}

// Launches kernels on a stream, or records them as a chain of kernel nodes of a graph under
// construction (when `graph` is set). Graphs are built node by node instead of by stream capture:
// a capture is invalidated by synchronising calls that other threads make meanwhile (stream and
// event synchronisation, copies from pageable memory), and the request loop's threads make such
// calls all the time.
struct Launcher {
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;
  hipError_t err = hipSuccess;

  template <class... P, class... A>
  void operator()(void (*f)(P...), dim3 grid, dim3 block, size_t shm, A&&... a) {
    if (!graph) {
      hipLaunchKernelGGL(f, grid, block, shm, stream, std::forward<A>(a)...);
      return;
    }
    if (err != hipSuccess) return;
    std::tuple<std::decay_t<P>...> values(std::forward<A>(a)...);
    void* params[sizeof...(P) + 1];
    std::apply(
        [&](auto&... e) {
          int k = 0;
          ((params[k++] = (void*)&e), ...);
        },
        values);
    hipKernelNodeParams kp{};
    kp.func = reinterpret_cast<void*>(f);
    kp.gridDim = grid;
    kp.blockDim = block;
    kp.sharedMemBytes = (unsigned)shm;
    kp.kernelParams = params;
    kp.extra = nullptr;
    hipGraphNode_t node = nullptr;
    err = hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &kp);
    if (err == hipSuccess) last = node;
  }
  // the first failure: a node that could not be added, or (on a stream) the last launch error
  hipError_t status() { return graph ? err : hipGetLastError(); }
};

// Launch-shape choices of the autotuner, per (parameter, nx, ny, precision, grids) key; -1 when
// unset (the caller then uses its own default). Only parameters that leave every result bit
// unchanged are tuned.
enum TuneParam {
  TUNE_DST_ROWS = 0,
  TUNE_LBM_LIST_BLOCKS = 1,
  TUNE_LBM_FUSE = 2,
  TUNE_LBM_ROWS = 3,
  TUNE_DST_PPT = 4,
  TUNE_LBM_NT = 5,
  TUNE_COUNT
};
int tune_get(int param, int nx, int ny, int precision, int grids);
void tune_set(int param, int nx, int ny, int precision, int grids, int value);

// Thread-local last error message returned by fb_last_error().
void set_error(const char* fmt, ...);
void clear_error();

// Per-block reduction slots: a reduction kernel writes one value per block and quantity into
// partials[q * kPartialBlocks + block]; finish_partials() folds them into the destination keys.
// This keeps large grids off a single address (same-address atomics serialise at one channel).
constexpr int kPartialBlocks = 8192;
constexpr int kPartialKeys = 16;

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  int xcd_count = 8;
  unsigned long long* partials = nullptr;  // kPartialKeys x kPartialBlocks, used in stream order
};

// What finish_partials() does with each quantity: fold it by sum (counts) or by max (keys
// from dkey()) and merge it into keys[dest] with one atomic.
struct FinishSpec {
  int n = 0;
  int dest[kPartialKeys];
  int sum[kPartialKeys];
  void add(int d, bool is_sum) {
    dest[n] = d;
    sum[n] = is_sum ? 1 : 0;
    ++n;
  }
};

// Enqueues the fold of `nblocks` per-block values of each quantity of `f` into `keys`.
int finish_partials(Ctx* c, unsigned nblocks, const FinishSpec& f, unsigned long long* keys);

// Grid of a grid-stride reduction kernel of 256-thread blocks over n items.
inline unsigned reduction_blocks(size_t n) {
  size_t b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > (size_t)kPartialBlocks ? (size_t)kPartialBlocks : b));
}

}  // namespace fb

using fb::dev_alloc;
using fb::dev_free;
using fb::host_alloc;
using fb::host_free;
using fb::graph_exec_free;
using fb::stream_create;
using fb::stream_free;

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

// Max that propagates NaN like NumPy's max (fmax would drop it).
__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

// Key of a magnitude for max reductions: NaN maps above +inf, so a max over keys propagates it.
__device__ __forceinline__ unsigned long long mag_key(double v) { return dkey(fabs(v)); }

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block reductions over a 256-thread block; the result is valid in thread 0. `lds` holds 4 values.
__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v, unsigned long long* lds) {
  v = wave_max_u64(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long r = v;
  if (threadIdx.x == 0) {
    r = lds[0] > lds[1] ? lds[0] : lds[1];
    r = r > lds[2] ? r : lds[2];
    r = r > lds[3] ? r : lds[3];
  }
  __syncthreads();
  return r;
}
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long* lds) {
  v = wave_sum_u64(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long r = v;
  if (threadIdx.x == 0) r = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return r;
}

// Block-reduce one quantity and store it in the block's partial slot q.
__device__ __forceinline__ void put_max(unsigned long long* part, int q, unsigned long long v,
                                        unsigned long long* lds) {
  v = block_max_u64(v, lds);
  if (threadIdx.x == 0) part[(size_t)q * kPartialBlocks + blockIdx.x] = v;
}
__device__ __forceinline__ void put_sum(unsigned long long* part, int q, unsigned long long v,
                                        unsigned long long* lds) {
  v = block_sum_u64(v, lds);
  if (threadIdx.x == 0) part[(size_t)q * kPartialBlocks + blockIdx.x] = v;
}

// Adds 1 to count[c] for every active lane of the wave. Neighbouring lanes with the same c
// share one atomic: particles are seeded in cell order and keep most of that order.
// Every lane of the wave must call it.
__device__ __forceinline__ void wave_run_add(int* count, int c, bool active) {
  const int lane = threadIdx.x & 63;
  const int cc = active ? c : -1;
  const int prev = __shfl_up(cc, 1, 64);
  const bool head = lane == 0 || prev != cc;
  const unsigned long long heads = __ballot(head);
  if (head && cc >= 0) {
    const unsigned long long rest = lane == 63 ? 0ull : heads >> (lane + 1);
    const int len = rest ? __ffsll((long long)rest) : 64 - lane;
    atomicAdd(&count[cc], len);
  }
}

// Claims a slot in cursor[c] for every active lane; lanes of one run get consecutive slots in
// lane order. Every lane of the wave must call it. Returns the claimed slot (or -1).
__device__ __forceinline__ int wave_run_claim(int* cursor, int c, bool active) {
  const int lane = threadIdx.x & 63;
  const int cc = active ? c : -1;
  const int prev = __shfl_up(cc, 1, 64);
  const bool head = lane == 0 || prev != cc;
  const unsigned long long heads = __ballot(head);
  int base = 0;
  if (head && cc >= 0) {
    const unsigned long long rest = lane == 63 ? 0ull : heads >> (lane + 1);
    const int len = rest ? __ffsll((long long)rest) : 64 - lane;
    base = atomicAdd(&cursor[cc], len);
  }
  const unsigned long long upto = lane == 63 ? heads : (heads & ((2ull << lane) - 1ull));
  const int hl = 63 - __clzll((long long)upto);
  base = __shfl(base, hl, 64);
  return cc >= 0 ? base + (lane - hl) : -1;
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
