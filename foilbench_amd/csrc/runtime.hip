// Context, error state and device selection of libfoilbench_amd.
#include <cstdarg>

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

}  // namespace fb

extern "C" {

const char* fb_last_error(void) { return fb::g_error.c_str(); }

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
  FB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  *out = reinterpret_cast<fb_ctx*>(c);
  return FB_OK;
}

int fb_ctx_destroy(fb_ctx* ctx) {
  if (!ctx) return FB_OK;
  auto* c = reinterpret_cast<fb::Ctx*>(ctx);
  FB_HIP(hipSetDevice(c->device));
  FB_HIP(hipStreamSynchronize(c->stream));
  FB_HIP(hipStreamDestroy(c->stream));
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
