#include "runtime.hpp"

#include <hip/hip_runtime_api.h>

namespace mscx {

static thread_local std::string g_last_error;

void set_last_error(const std::string &m) { g_last_error = m; }

Runtime &runtime() {
  static Runtime rt;
  return rt;
}

void ensure_init() {
  Runtime &rt = runtime();
  if (rt.device >= 0) {
    MSCX_HIP(hipSetDevice(rt.device));
    return;
  }
  int n = 0;
  MSCX_HIP(hipGetDeviceCount(&n));
  MSCX_REQUIRE(n > 0, MSCX_EDEVICE, "no HIP device is visible");
  MSCX_HIP(hipSetDevice(0));
  MSCX_HIP(hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking));
  rt.device = 0;
}

void *Scratch::reserve_dev(size_t n) {
  if (n > dev_bytes) {
    if (dev) MSCX_HIP(hipFree(dev));
    dev = nullptr;
    size_t want = align_up(n + n / 2, 1 << 16);
    MSCX_HIP(hipMalloc(&dev, want));
    dev_bytes = want;
  }
  return dev;
}

void *Scratch::reserve_host(size_t n) {
  if (n > host_bytes) {
    if (host) MSCX_HIP(hipHostFree(host));
    host = nullptr;
    size_t want = align_up(n + n / 2, 1 << 16);
    MSCX_HIP(hipHostMalloc(&host, want, hipHostMallocDefault));
    host_bytes = want;
  }
  return host;
}

Scratch::~Scratch() {
  // Buffers live until process exit; the runtime may already be torn down
  // when static destructors run, so nothing is released here.
}

}  // namespace mscx

extern "C" {

const char *mscx_last_error(void) { return mscx::g_last_error.c_str(); }

int mscx_version(void) { return 1; }

int mscx_init(int device) {
  return mscx::guarded([&] {
    mscx::Runtime &rt = mscx::runtime();
    std::lock_guard<std::mutex> lk(rt.mu);
    int n = 0;
    MSCX_HIP(hipGetDeviceCount(&n));
    MSCX_REQUIRE(device >= 0 && device < n, MSCX_EINVAL, "device index out of range");
    if (rt.device == device) return;
    MSCX_REQUIRE(rt.device < 0, MSCX_ESTATE, "mscx_init: already initialised on another device");
    MSCX_HIP(hipSetDevice(device));
    MSCX_HIP(hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking));
    rt.device = device;
  });
}

int mscx_shutdown(void) {
  return mscx::guarded([&] {
    mscx::Runtime &rt = mscx::runtime();
    std::lock_guard<std::mutex> lk(rt.mu);
    if (rt.device < 0) return;
    MSCX_HIP(hipStreamSynchronize(rt.stream));
  });
}

}  // extern "C"
