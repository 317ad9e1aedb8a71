// Seeded synthetic scalar fields (raw float32, x fastest), the inputs of the
// benchmark and test configurations.  Every sample is a pure function of
// (seed, coordinates), so generation is parallel over rows and bit-for-bit
// reproducible on any x86-64 host.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "runtime.hpp"

namespace mscx {

static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }
static inline uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  return mix64(mix64(mix64(seed ^ mix64(a)) ^ b) ^ c);
}

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() { return mix64(s++); }
  double uniform() { return u01(next()); }
  double normal() {  // Box-Muller
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

template <class F>
static void parallel_rows(int sy, F f) {
  int nt = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
  if ((int64_t)sy < 64) nt = 1;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([=] {
      for (int y = t; y < sy; y += nt) f(y);
    });
  for (auto &x : th) x.join();
}

static void check_dims(int sx, int sy) {
  MSCX_REQUIRE(sx >= 2 && sy >= 2 && sx <= 16384 && sy <= 16384, MSCX_EINVAL, "grid sides must be in [2, 16384]");
}

}  // namespace mscx

using namespace mscx;

extern "C" {

int mscx_gen_gaussians(float *out, int32_t sx, int32_t sy, int32_t n, double noise, uint64_t seed) {
  return guarded([&] {
    MSCX_REQUIRE(out && n >= 0, MSCX_EINVAL, "bad argument");
    check_dims(sx, sy);
    Rng r(seed);
    std::vector<double> cx(n), cy(n), sg(n), a(n);
    for (int k = 0; k < n; ++k) {
      cx[k] = r.uniform() * sx;
      cy[k] = r.uniform() * sy;
      sg[k] = 8.0 + 56.0 * r.uniform();
      a[k] = r.normal();
    }
    // separable factors: exp(-(dx^2+dy^2)/2s^2) = ex(x) * ey(y)
    std::vector<double> ex((size_t)n * sx), ey((size_t)n * sy);
    for (int k = 0; k < n; ++k) {
      double inv = 1.0 / (2.0 * sg[k] * sg[k]);
      for (int x = 0; x < sx; ++x) ex[(size_t)k * sx + x] = std::exp(-(x - cx[k]) * (x - cx[k]) * inv);
      for (int y = 0; y < sy; ++y) ey[(size_t)k * sy + y] = a[k] * std::exp(-(y - cy[k]) * (y - cy[k]) * inv);
    }
    uint64_t nseed = mix64(seed ^ 0x6E6F697365ull);
    parallel_rows(sy, [&](int y) {
      float *row = out + (size_t)y * sx;
      for (int x = 0; x < sx; ++x) {
        double v = 0.0;
        for (int k = 0; k < n; ++k) v += ey[(size_t)k * sy + y] * ex[(size_t)k * sx + x];
        v += noise * (2.0 * u01(hash3(nseed, 0, (uint64_t)x, (uint64_t)y)) - 1.0);
        row[x] = (float)v;
      }
    });
  });
}

int mscx_gen_value_noise(float *out, int32_t sx, int32_t sy, int32_t octaves, double lac, double gain,
                         double period, uint64_t seed) {
  return guarded([&] {
    MSCX_REQUIRE(out && octaves >= 1 && octaves <= 32 && lac > 0 && period > 0, MSCX_EINVAL, "bad argument");
    check_dims(sx, sy);
    parallel_rows(sy, [&](int y) {
      float *row = out + (size_t)y * sx;
      for (int x = 0; x < sx; ++x) {
        double v = 0.0, amp = 1.0, per = period;
        for (int o = 0; o < octaves; ++o) {
          double u = x / per, w = y / per;
          double fi = std::floor(u), fj = std::floor(w);
          double fu = u - fi, fw = w - fj;
          double su = fu * fu * (3.0 - 2.0 * fu), sw = fw * fw * (3.0 - 2.0 * fw);
          int64_t i = (int64_t)fi, j = (int64_t)fj;
          uint64_t os = mix64(seed + 0x1000003ull * (uint64_t)o);
          double v00 = u01(hash3(os, 1, (uint64_t)i, (uint64_t)j));
          double v10 = u01(hash3(os, 1, (uint64_t)(i + 1), (uint64_t)j));
          double v01 = u01(hash3(os, 1, (uint64_t)i, (uint64_t)(j + 1)));
          double v11 = u01(hash3(os, 1, (uint64_t)(i + 1), (uint64_t)(j + 1)));
          double a0 = v00 + su * (v10 - v00), a1 = v01 + su * (v11 - v01);
          v += amp * (a0 + sw * (a1 - a0));
          amp *= gain;
          per /= lac;
        }
        row[x] = (float)v;
      }
    });
  });
}

int mscx_gen_white(float *out, int32_t sx, int32_t sy, int32_t levels, uint64_t seed) {
  return guarded([&] {
    MSCX_REQUIRE(out && levels >= 0, MSCX_EINVAL, "bad argument");
    check_dims(sx, sy);
    parallel_rows(sy, [&](int y) {
      float *row = out + (size_t)y * sx;
      for (int x = 0; x < sx; ++x) {
        double u = u01(hash3(seed, 2, (uint64_t)x, (uint64_t)y));
        row[x] = levels > 0 ? (float)(std::floor(u * levels) / levels) : (float)u;
      }
    });
  });
}

}  // extern "C"
