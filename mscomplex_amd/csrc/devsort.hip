// Device scan and stable radix sort (see devsort.hpp).
#include <hip/hip_runtime.h>

#include "devsort.hpp"
#include "runtime.hpp"

namespace mscx {

namespace {

constexpr int SCAN_PER = 8, SCAN_BLK = 256 * SCAN_PER;
constexpr int RS_ITEMS = 16, RS_BLK = 256 * RS_ITEMS;  // pairs per sort block

// 256-thread exclusive scan of one value per thread (wave64 shuffles).
__device__ uint32_t block_scan(uint32_t v, uint32_t &total) {
  __shared__ uint32_t sw[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sw[w] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q < w) base += sw[q];
    tot += sw[q];
  }
  __syncthreads();
  total = tot;
  return base + x - v;
}

__global__ __launch_bounds__(256) void k_scan_sum(const uint32_t *__restrict__ in, uint64_t n,
                                                  uint32_t *__restrict__ part) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_BLK + (uint64_t)threadIdx.x * SCAN_PER;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t tot;
  block_scan(s, tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// in and out may be the same array
__global__ __launch_bounds__(256) void k_scan_down(const uint32_t *in, uint32_t *out, uint64_t n,
                                                   const uint32_t *__restrict__ partex, uint32_t *total) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_BLK + (uint64_t)threadIdx.x * SCAN_PER;
  uint32_t v[SCAN_PER];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t ex = block_scan(s, tot) + (partex ? partex[blockIdx.x] : 0u);
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (total && threadIdx.x == 0) *total = tot;
}

// The lanes of this wave whose valid digit equals this lane's (eight
// ballots, one per digit bit).
__device__ __forceinline__ uint64_t digit_peers(bool valid, uint32_t d) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int bit = 0; bit < 8; ++bit) {
    const bool on = (d >> bit) & 1u;
    const uint64_t b = __ballot(valid && on);
    peers &= on ? b : ~b;
  }
  return peers;
}

// Digit histogram of one block of RS_BLK pairs, stored digit-major:
// hist[d * nblocks + b].  Equal digits of a wave add once (keys often
// repeat in long runs, which would serialise per-element LDS atomics).
__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t *__restrict__ keys, uint64_t n, int shift,
                                                 uint32_t nblocks, uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[256];
  const int t = threadIdx.x, lane = t & 63;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  h[t] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_BLK;
  for (int r = 0; r < RS_ITEMS; ++r) {
    const uint64_t i = base + (uint64_t)r * 256 + t;
    const bool valid = i < n;
    const uint32_t d = valid ? (keys[i] >> shift) & 255u : 0u;
    const uint64_t peers = digit_peers(valid, d);
    if (valid && !(peers & below)) atomicAdd(&h[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  hist[(uint64_t)t * nblocks + blockIdx.x] = h[t];
}

// Stable scatter of one block: its pairs in input order (round r, then
// wave, then lane) go to their digit's next positions.  Within a wave the
// lanes of equal digits find each other with eight ballots (one per digit
// bit); a lane's rank is the number of those peers below it, and the
// lowest peer publishes the wave's count of that digit.
__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t *__restrict__ keys,
                                                    const uint32_t *__restrict__ vals, uint32_t *__restrict__ ko,
                                                    uint32_t *__restrict__ vo, uint64_t n, int shift,
                                                    uint32_t nblocks, const uint32_t *__restrict__ off) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[4][256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  run[t] = off[(uint64_t)t * nblocks + blockIdx.x];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint64_t base = (uint64_t)blockIdx.x * RS_BLK;
  for (int r = 0; r < RS_ITEMS; ++r) {
#pragma unroll
    for (int q = 0; q < 4; ++q) wc[q][t] = 0;
    __syncthreads();
    const uint64_t i = base + (uint64_t)r * 256 + t;
    const bool valid = i < n;
    const uint32_t k = valid ? keys[i] : 0u, v = valid ? vals[i] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    const uint64_t peers = digit_peers(valid, d);
    const uint32_t rank = (uint32_t)__popcll(peers & below);
    if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int q = 0; q < w; ++q) pos += wc[q][d];
      ko[pos] = k;
      vo[pos] = v;
    }
    __syncthreads();
    run[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
  }
}

}  // namespace

uint64_t scan_temp_words(uint64_t n) {
  const uint64_t nb = (n + SCAN_BLK - 1) / SCAN_BLK;
  return nb > 1 ? nb + scan_temp_words(nb) : 0;
}

void dev_scan(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *d_total, uint32_t *tmp, hipStream_t s) {
  if (n == 0) {
    if (d_total) MSCX_HIP(hipMemsetAsync(d_total, 0, 4, s));
    return;
  }
  const uint64_t nb = (n + SCAN_BLK - 1) / SCAN_BLK;
  if (nb == 1) {
    k_scan_down<<<1, 256, 0, s>>>(in, out, n, nullptr, d_total);
    return;
  }
  uint32_t *part = tmp;
  k_scan_sum<<<(unsigned)nb, 256, 0, s>>>(in, n, part);
  dev_scan(part, part, nb, d_total, tmp + nb, s);
  k_scan_down<<<(unsigned)nb, 256, 0, s>>>(in, out, n, part, nullptr);
}

uint64_t sort_temp_words(uint64_t n) {
  const uint64_t nb = (n + RS_BLK - 1) / RS_BLK;
  return 256 * nb + scan_temp_words(256 * nb);
}

bool radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys2, uint32_t *vals2, uint64_t n, int bits,
                      uint32_t *tmp, hipStream_t s) {
  if (n <= 1 || bits <= 0) return false;
  const uint64_t nb = (n + RS_BLK - 1) / RS_BLK;
  MSCX_REQUIRE(nb < (1ull << 31), MSCX_EINVAL, "too many pairs to sort");
  uint32_t *hist = tmp, *stmp = tmp + 256 * nb;
  bool alt = false;
  for (int shift = 0; shift < bits; shift += 8) {
    const uint32_t *ki = alt ? keys2 : keys, *vi = alt ? vals2 : vals;
    uint32_t *kout = alt ? keys : keys2, *vout = alt ? vals : vals2;
    k_rs_hist<<<(unsigned)nb, 256, 0, s>>>(ki, n, shift, (uint32_t)nb, hist);
    dev_scan(hist, hist, 256 * nb, nullptr, stmp, s);
    k_rs_scatter<<<(unsigned)nb, 256, 0, s>>>(ki, vi, kout, vout, n, shift, (uint32_t)nb, hist);
    alt = !alt;
  }
  return alt;
}

}  // namespace mscx
