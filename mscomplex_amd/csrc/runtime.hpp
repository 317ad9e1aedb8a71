// Error handling and device runtime shared by the C-ABI implementation.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <exception>
#include <thread>
#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mscx.h"

namespace mscx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &m);

#define MSCX_HIP(x)                                                                                       \
  do {                                                                                                    \
    hipError_t e_ = (x);                                                                                  \
    if (e_ != hipSuccess)                                                                                 \
      throw ::mscx::Error(MSCX_EDEVICE, std::string(#x) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

#define MSCX_REQUIRE(cond, code, msg) \
  do {                                \
    if (!(cond)) throw ::mscx::Error((code), (msg)); \
  } while (0)

// Runs f, converting exceptions to a status code and the thread's last error.
template <class F>
int guarded(F &&f) {
  try {
    f();
    return MSCX_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return MSCX_ENOMEM;
  } catch (const std::logic_error &e) {
    set_last_error(e.what());
    return MSCX_ELOGIC;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return MSCX_EDEVICE;
  }
}

// Device state: the selected device, the work stream and a mutex that
// serialises batched launches issued from several host threads.
struct Runtime {
  int device = -1;
  hipStream_t stream = nullptr;  // tile stages (run_tiles), serialised by mu
  hipStream_t aux = nullptr;     // manifold passes (label kernel and their copies)
  std::mutex mu;
  // stage graphs (run_tiles), guarded by mu
  bool graphs = true;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uintptr_t> graph_key;  // launch configuration graph_exec was built for
  uint64_t graph_builds = 0, graph_replays = 0;
};
Runtime &runtime();
void ensure_init();

// Host threads the parallel host phases use: min(hardware threads, 16), or
// the calling thread's cap when a ThreadCap is live on it (the file run's
// num_parallel; 1 runs every host phase on the caller).  parallel loops pass
// their caller's cap on to their workers.
int host_threads();
int thread_cap();
void set_thread_cap(int n);
struct ThreadCap {
  int prev;
  explicit ThreadCap(int n) : prev(thread_cap()) {
    if (n > 0) set_thread_cap(n);
  }
  ~ThreadCap() { set_thread_cap(prev); }
  ThreadCap(const ThreadCap &) = delete;
  ThreadCap &operator=(const ThreadCap &) = delete;
};

// Dynamic parallel loop over [0, n) on up to `nt` threads.  When items
// fail, the exception of the lowest failing index is rethrown after all
// threads have joined: the failure a sequential loop would have met first,
// whatever the threads' timing (items after a failure still run; callers
// discard their results).
template <class F>
void parallel_for(int n, F f, int nt = 0) {
  if (nt <= 0) nt = host_threads();
  nt = std::min(nt, n);
  if (n <= 1 || nt <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> errs(nt);
  std::vector<int> err_at(nt, n);
  const int cap = thread_cap();
  auto body = [&](int t) {
    ThreadCap keep(cap);
    for (int i = next++; i < n; i = next++) {
      try {
        f(i);
      } catch (...) {
        if (i < err_at[t]) {
          err_at[t] = i;
          errs[t] = std::current_exception();
        }
      }
    }
  };
  // worker 0 is the calling thread; a worker that cannot be started leaves
  // its share to the others
  for (int t = 1; t < nt; ++t) {
    try {
      th.emplace_back(body, t);
    } catch (...) {
      break;
    }
  }
  body(0);
  for (auto &x : th) x.join();
  int first = -1;
  for (int t = 0; t < nt; ++t)
    if (errs[t] && (first < 0 || err_at[t] < err_at[first])) first = t;
  if (first >= 0) std::rethrow_exception(errs[first]);
}

// Growable device scratch and pinned host staging used by batched launches.
struct Scratch {
  void *dev = nullptr;
  size_t dev_bytes = 0;
  void *host = nullptr;
  size_t host_bytes = 0;
  void *reserve_dev(size_t n);
  void *reserve_host(size_t n);
  ~Scratch();
};

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Caching device allocator: freed blocks are kept per size class and reused,
// so steady-state runs make no hipMalloc/hipFree calls.  in_use() is the
// HBM the library currently holds for live objects (the serving loop admits
// work against it).
class DevicePool {
 public:
  void *alloc(size_t bytes);
  void free(void *p);
  size_t in_use() const { return in_use_; }
  size_t cached() const { return cached_; }
  void trim();  // returns cached blocks to the runtime

 private:
  std::mutex mu_;
  std::multimap<size_t, void *> free_;  // class size -> block
  std::map<void *, size_t> live_;       // block -> class size
  size_t in_use_ = 0, cached_ = 0;
  static size_t size_class(size_t n);
};
DevicePool &device_pool();

// Optional per-stage device timing with HIP events recorded on the work
// stream around each stage's launches (enabled by mscx_profile).
// Stages 0-6 are device stages timed with HIP events; the manifold pass's
// host parts are wall-clock times of the calls (summed over concurrent
// calls): 7 critical-point pass, 8 label tables, 10 the disc tracing on the
// GPU with its transfers and scan, 13 the extremum discs' grouping with the
// lists' copy into the complexes; 9, 11 and 12 are host-thread times
// (summed over the threads) of waiting for the labels, of the lists' copy
// (or host tracing) and of grouping; 14 the codes' download to the host;
// 15 the host copies of the leaves' complexes after a merge tree on the GPU.
enum Stage {
  ST_GRADIENT = 0, ST_SCAN, ST_COMPACT, ST_TRACE, ST_OWNER, ST_INCIDENCE, ST_HOST_COMPLEX,
  ST_DISC_PREP, ST_DISC_TABLES, ST_DISC_LABELS, ST_DISC_PATHS, ST_DISC_TRACE_CPU, ST_DISC_GROUP_CPU, ST_DISC_GROUP,
  ST_DISC_CODES, ST_LEAF_COPIES, ST_COUNT
};
struct StageTimer {
  bool on = false;
  hipEvent_t a[ST_COUNT] = {}, b[ST_COUNT] = {};
  bool pending[ST_COUNT] = {};
  double ms[ST_COUNT] = {};
  uint32_t count[ST_COUNT] = {};
  void ensure(int k);  // creates stage k's event pair
  void begin(int k, hipStream_t s);
  void end(int k, hipStream_t s);
  void collect();  // call after a stream synchronisation
  void add_host(int k, double seconds);  // a host stage's wall-clock time
};
StageTimer &stage_timer();

}  // namespace mscx
