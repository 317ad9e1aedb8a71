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

StageTimer &stage_timer() {
  static StageTimer t;
  return t;
}

void StageTimer::begin(int k, hipStream_t s) {
  if (!on) return;
  if (!a[k]) {
    MSCX_HIP(hipEventCreate(&a[k]));
    MSCX_HIP(hipEventCreate(&b[k]));
  }
  MSCX_HIP(hipEventRecord(a[k], s));
}

void StageTimer::end(int k, hipStream_t s) {
  if (!on) return;
  MSCX_HIP(hipEventRecord(b[k], s));
  pending[k] = true;
}

void StageTimer::collect() {
  if (!on) return;
  for (int k = 0; k < ST_COUNT; ++k) {
    if (!pending[k]) continue;
    float m = 0;
    MSCX_HIP(hipEventElapsedTime(&m, a[k], b[k]));
    ms[k] += m;
    count[k]++;
    pending[k] = false;
  }
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

int mscx_profile(int enable) {
  return mscx::guarded([&] {
    mscx::StageTimer &t = mscx::stage_timer();
    t.on = enable != 0;
    for (int k = 0; k < mscx::ST_COUNT; ++k) {
      t.ms[k] = 0;
      t.count[k] = 0;
      t.pending[k] = false;
    }
  });
}

int mscx_profile_read(double *ms, uint32_t *counts, int n) {
  return mscx::guarded([&] {
    mscx::StageTimer &t = mscx::stage_timer();
    for (int k = 0; k < n && k < mscx::ST_COUNT; ++k) {
      if (ms) ms[k] = t.ms[k];
      if (counts) counts[k] = t.count[k];
    }
  });
}

int mscx_dev_alloc(size_t bytes, void **out) {
  return mscx::guarded([&] {
    MSCX_REQUIRE(out, MSCX_EINVAL, "null argument");
    mscx::ensure_init();
    *out = nullptr;
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    MSCX_REQUIRE(e == hipSuccess, MSCX_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  });
}

int mscx_dev_free(void *p) {
  return mscx::guarded([&] {
    if (!p) return;
    MSCX_HIP(hipStreamSynchronize(mscx::runtime().stream));
    MSCX_HIP(hipFree(p));
  });
}

int mscx_dev_upload(void *dst, const void *src, size_t bytes) {
  return mscx::guarded([&] {
    mscx::ensure_init();
    MSCX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}

int mscx_dev_download(void *dst, const void *src, size_t bytes) {
  return mscx::guarded([&] {
    mscx::ensure_init();
    MSCX_HIP(hipStreamSynchronize(mscx::runtime().stream));
    MSCX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
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
