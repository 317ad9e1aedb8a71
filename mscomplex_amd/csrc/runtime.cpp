#include "runtime.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <thread>

namespace mscx {

static thread_local std::string g_last_error;

void set_last_error(const std::string &m) { g_last_error = m; }

static thread_local int t_thread_cap = 0;

int thread_cap() { return t_thread_cap; }
void set_thread_cap(int n) { t_thread_cap = n; }

int host_threads() {
  if (t_thread_cap > 0) return t_thread_cap;
  return (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
}

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
  static std::mutex init_mu;  // first use may come from several threads
  std::lock_guard<std::mutex> lk(init_mu);
  if (rt.device >= 0) {
    MSCX_HIP(hipSetDevice(rt.device));
    return;
  }
  int n = 0;
  MSCX_HIP(hipGetDeviceCount(&n));
  MSCX_REQUIRE(n > 0, MSCX_EDEVICE, "no HIP device is visible");
  MSCX_HIP(hipSetDevice(0));
  MSCX_HIP(hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking));
  MSCX_HIP(hipStreamCreateWithFlags(&rt.aux, hipStreamNonBlocking));
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

size_t DevicePool::size_class(size_t n) {
  if (n <= 4096) return 4096;
  // classes at 1/8 steps of each power of two
  size_t p = 1;
  while (p * 2 <= n) p *= 2;
  size_t step = p / 8;
  return align_up(n, step);
}

void *DevicePool::alloc(size_t bytes) {
  size_t cls = size_class(bytes);
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = free_.find(cls);
    if (it != free_.end()) {
      void *p = it->second;
      free_.erase(it);
      cached_ -= cls;
      live_[p] = cls;
      in_use_ += cls;
      return p;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, cls);
  if (e != hipSuccess) {
    trim();  // give cached blocks back and retry once
    e = hipMalloc(&p, cls);
  }
  MSCX_REQUIRE(e == hipSuccess, MSCX_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(mu_);
  live_[p] = cls;
  in_use_ += cls;
  return p;
}

void DevicePool::free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  size_t cls = it->second;
  live_.erase(it);
  in_use_ -= cls;
  free_.emplace(cls, p);
  cached_ += cls;
}

void DevicePool::trim() {
  std::multimap<size_t, void *> f;
  {
    std::lock_guard<std::mutex> lk(mu_);
    f.swap(free_);
    cached_ = 0;
  }
  for (auto &kv : f) (void)hipFree(kv.second);
}

DevicePool &device_pool() {
  static DevicePool *p = new DevicePool;  // lives until process exit
  return *p;
}

StageTimer &stage_timer() {
  static StageTimer t;
  return t;
}

void StageTimer::ensure(int k) {
  if (!a[k]) {
    MSCX_HIP(hipEventCreate(&a[k]));
    MSCX_HIP(hipEventCreate(&b[k]));
  }
}

void StageTimer::begin(int k, hipStream_t s) {
  if (!on) return;
  ensure(k);
  MSCX_HIP(hipEventRecord(a[k], s));
}

void StageTimer::end(int k, hipStream_t s) {
  if (!on) return;
  MSCX_HIP(hipEventRecord(b[k], s));
  pending[k] = true;
}

void StageTimer::add_host(int k, double seconds) {
  if (!on) return;
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  ms[k] += 1e3 * seconds;
  count[k]++;
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
    MSCX_HIP(hipStreamCreateWithFlags(&rt.aux, hipStreamNonBlocking));
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

int mscx_set_graphs(int on, int *previous) {
  return mscx::guarded([&] {
    mscx::Runtime &rt = mscx::runtime();
    std::lock_guard<std::mutex> lk(rt.mu);
    if (previous) *previous = rt.graphs ? 1 : 0;
    rt.graphs = on != 0;
  });
}

int mscx_graph_stats(uint64_t *builds, uint64_t *replays) {
  return mscx::guarded([&] {
    mscx::Runtime &rt = mscx::runtime();
    std::lock_guard<std::mutex> lk(rt.mu);
    if (builds) *builds = rt.graph_builds;
    if (replays) *replays = rt.graph_replays;
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
    *out = mscx::device_pool().alloc(bytes ? bytes : 1);
  });
}

int mscx_dev_free(void *p) {
  return mscx::guarded([&] {
    if (!p) return;
    MSCX_HIP(hipStreamSynchronize(mscx::runtime().stream));
    mscx::device_pool().free(p);
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
    if (rt.graph_exec) {
      MSCX_HIP(hipGraphExecDestroy(rt.graph_exec));
      rt.graph_exec = nullptr;
      rt.graph_key.clear();
    }
  });
}

}  // extern "C"
