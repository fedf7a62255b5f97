// Context, error state and device selection of libfoilbench_amd.
#include <algorithm>
#include <cstdarg>
#include <map>
#include <tuple>
#include <unordered_map>

#include "fb_common.h"

namespace fb {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
}

void clear_error() { g_error.clear(); }

std::recursive_mutex& hip_lock() {
  static std::recursive_mutex m;
  return m;
}

namespace {
struct TuneKey {
  int param, nx, ny, prec, grids;
  bool operator<(const TuneKey& o) const {
    return std::tie(param, nx, ny, prec, grids) < std::tie(o.param, o.nx, o.ny, o.prec, o.grids);
  }
};
std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}
std::map<TuneKey, int>& tune_table() {
  static std::map<TuneKey, int> t;
  return t;
}
}  // namespace

int tune_get(int param, int nx, int ny, int precision, int grids) {
  std::lock_guard<std::mutex> g(tune_mutex());
  auto& t = tune_table();
  auto it = t.find({param, nx, ny, precision, grids});
  return it == t.end() ? -1 : it->second;
}

void tune_set(int param, int nx, int ny, int precision, int grids, int value) {
  std::lock_guard<std::mutex> g(tune_mutex());
  if (value < 0) tune_table().erase({param, nx, ny, precision, grids});
  else tune_table()[{param, nx, ny, precision, grids}] = value;
}

namespace {
struct BlockCache {
  std::unordered_map<void*, std::pair<int, size_t>> live;     // block -> (device, size class)
  std::multimap<std::pair<int, size_t>, void*> idle;          // (device, size class) -> free block
  size_t idle_bytes = 0;
};
BlockCache& cache() {
  static BlockCache* c = new BlockCache();  // never destroyed: blocks may be freed at exit
  return *c;
}
// Size classes: powers of two up to 1 MiB, then whole MiB.
size_t size_class(size_t b) {
  if (b <= 512) return 512;
  if (b <= (size_t(1) << 20)) {
    size_t c = 512;
    while (c < b) c <<= 1;
    return c;
  }
  const size_t mib = size_t(1) << 20;
  return (b + mib - 1) / mib * mib;
}
constexpr size_t kMaxIdleBytes = size_t(32) << 30;
}  // namespace

void release_cached_blocks() {
  std::lock_guard<std::recursive_mutex> g(hip_lock());
  BlockCache& c = cache();
  for (auto& kv : c.idle) (void)hipFree(kv.second);
  c.idle.clear();
  c.idle_bytes = 0;
}

hipError_t cached_alloc(void** p, size_t bytes) {
  std::lock_guard<std::recursive_mutex> g(hip_lock());
  BlockCache& c = cache();
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const size_t sz = size_class(bytes);
  auto it = c.idle.find({dev, sz});
  if (it != c.idle.end()) {
    *p = it->second;
    c.idle.erase(it);
    c.idle_bytes -= sz;
    c.live[*p] = {dev, sz};
    return hipSuccess;
  }
  e = hipMalloc(p, sz);
  if (e == hipErrorOutOfMemory && c.idle_bytes > 0) {
    (void)hipGetLastError();
    release_cached_blocks();
    e = hipMalloc(p, sz);
  }
  if (e == hipSuccess) c.live[*p] = {dev, sz};
  return e;
}

hipError_t cached_free(void* p) {
  if (!p) return hipSuccess;
  std::lock_guard<std::recursive_mutex> g(hip_lock());
  BlockCache& c = cache();
  auto it = c.live.find(p);
  if (it == c.live.end()) return hipFree(p);
  const auto key = it->second;
  c.live.erase(it);
  if (c.idle_bytes + key.second > kMaxIdleBytes) return hipFree(p);
  c.idle.insert({key, p});
  c.idle_bytes += key.second;
  return hipSuccess;
}

// Streaming copy used to calibrate HBM counters and the achievable bandwidth.
__global__ void __launch_bounds__(256) copy_f4_kernel(size_t n4, const float4* __restrict__ a, float4* __restrict__ b) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  size_t stride = (size_t)gridDim.x * 256;
  for (; i < n4; i += stride) b[i] = a[i];
}

// Folds the per-block values of quantity blockIdx.x (fixed order; eight loads in flight per
// thread, since the slots were just written and come from memory).
__global__ void __launch_bounds__(256) finish_partials_kernel(const unsigned long long* __restrict__ part,
                                                              unsigned nb, FinishSpec f, unsigned long long* keys) {
  __shared__ unsigned long long lds[4];
  const int q = blockIdx.x;
  const unsigned long long* p = part + (size_t)q * kPartialBlocks;
  const bool sum = f.sum[q] != 0;
  unsigned long long v = 0;
  for (unsigned i0 = 0; i0 < nb; i0 += 8 * 256) {
    unsigned long long w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const unsigned i = i0 + u * 256 + threadIdx.x;
      w[u] = i < nb ? p[i] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) v = sum ? v + w[u] : (w[u] > v ? w[u] : v);
  }
  v = sum ? block_sum_u64(v, lds) : block_max_u64(v, lds);
  if (threadIdx.x == 0) {
    if (sum) {
      if (v) atomicAdd(&keys[f.dest[q]], v);
    } else {
      key_max(&keys[f.dest[q]], v);
    }
  }
}

int finish_partials(Ctx* c, unsigned nblocks, const FinishSpec& f, unsigned long long* keys) {
  FB_REQUIRE(nblocks <= (unsigned)kPartialBlocks && f.n <= kPartialKeys, FB_ERR_ARG,
             "partial reduction over %u blocks / %d quantities exceeds the slots", nblocks, f.n);
  if (f.n == 0) return FB_OK;
  hipLaunchKernelGGL(finish_partials_kernel, dim3(f.n), dim3(256), 0, c->stream, (const unsigned long long*)c->partials,
                     nblocks, f, keys);
  FB_HIP(hipGetLastError());
  return FB_OK;
}

}  // namespace fb

extern "C" {

int fb_copy_benchmark(fb_ctx* ctx, int64_t bytes, int reps, double* ms_per_copy) {
  FB_REQUIRE(ctx && ms_per_copy, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(bytes >= 16 && bytes % 16 == 0 && reps >= 1, FB_ERR_ARG, "bytes must be a positive multiple of 16");
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  void *a = nullptr, *b = nullptr;
  FB_HIP(dev_alloc(&a, (size_t)bytes));
  FB_HIP(dev_alloc(&b, (size_t)bytes));
  FB_HIP(hipMemsetAsync(a, 0, (size_t)bytes, c->stream));
  size_t n4 = (size_t)bytes / 16;
  unsigned blocks = (unsigned)std::min<size_t>((n4 + 255) / 256, (size_t)c->cu_count * 32);
  hipEvent_t e0, e1;
  FB_HIP(hipEventCreate(&e0));
  FB_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(fb::copy_f4_kernel, dim3(blocks), dim3(256), 0, c->stream, n4, (const float4*)a, (float4*)b);
  FB_HIP(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(fb::copy_f4_kernel, dim3(blocks), dim3(256), 0, c->stream, n4, (const float4*)a, (float4*)b);
  FB_HIP(hipEventRecord(e1, c->stream));
  FB_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  FB_HIP(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_copy = ms / reps;
  FB_HIP(hipEventDestroy(e0));
  FB_HIP(hipEventDestroy(e1));
  FB_HIP(dev_free(a));
  FB_HIP(dev_free(b));
  return FB_OK;
}

const char* fb_last_error(void) { return fb::g_error.c_str(); }

int fb_tune_set(int param, int nx, int ny, int precision, int grids, int value) {
  FB_REQUIRE(param >= 0 && param < fb::TUNE_COUNT, FB_ERR_ARG, "unknown tuning parameter %d", param);
  fb::tune_set(param, nx, ny, precision, grids, value);
  return FB_OK;
}

int fb_tune_get(int param, int nx, int ny, int precision, int grids, int* value) {
  FB_REQUIRE(value && param >= 0 && param < fb::TUNE_COUNT, FB_ERR_ARG, "bad tuning query");
  *value = fb::tune_get(param, nx, ny, precision, grids);
  return FB_OK;
}

int fb_version(void) { return 1; }

int fb_device_count(int* count) {
  FB_REQUIRE(count != nullptr, FB_ERR_ARG, "count is NULL");
  FB_HIP(hipGetDeviceCount(count));
  return FB_OK;
}

int fb_ctx_create(int device, fb_ctx** out) {
  FB_REQUIRE(out != nullptr, FB_ERR_ARG, "out is NULL");
  int n = 0;
  FB_HIP(hipGetDeviceCount(&n));
  FB_REQUIRE(device >= 0 && device < n, FB_ERR_ARG, "device %d out of range (%d devices)", device, n);
  FB_HIP(hipSetDevice(device));
  auto* c = new fb::Ctx();
  c->device = device;
  hipDeviceProp_t prop;
  FB_HIP(hipGetDeviceProperties(&prop, device));
  c->cu_count = prop.multiProcessorCount;
  // Default-priority stream, non-blocking with respect to the null stream.
  FB_HIP(stream_create(&c->stream));
  if (dev_alloc(&c->partials, (size_t)fb::kPartialKeys * fb::kPartialBlocks * sizeof(unsigned long long)) != hipSuccess) {
    stream_free(c->stream);
    delete c;
    fb::set_error("fb_ctx_create: cannot allocate the reduction slots");
    return FB_ERR_NOMEM;
  }
  *out = reinterpret_cast<fb_ctx*>(c);
  return FB_OK;
}

int fb_ctx_destroy(fb_ctx* ctx) {
  if (!ctx) return FB_OK;
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  FB_HIP(hipStreamSynchronize(c->stream));
  FB_HIP(dev_free(c->partials));
  FB_HIP(stream_free(c->stream));
  delete c;
  return FB_OK;
}

int fb_ctx_synchronize(fb_ctx* ctx) {
  FB_REQUIRE(ctx != nullptr, FB_ERR_ARG, "ctx is NULL");
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  FB_HIP(hipStreamSynchronize(c->stream));
  return FB_OK;
}

}  // extern "C"
