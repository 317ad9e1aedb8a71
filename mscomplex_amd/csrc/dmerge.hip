// Device-resident complexes and the merge tree on the GPU (see dmerge.hpp).
//
// Every kernel here is a sweep over nodes or list entries (integer work on
// HBM-resident arrays; no step is a contraction); scans are devsort.hip's.
// Host code drives the kernels on the caller's
// stream and synchronises only where a count decides an allocation and
// around the host pass over a shared line.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <stdexcept>
#include <thread>

#include "devsort.hpp"
#include "dmerge.hpp"
#include "runtime.hpp"
#include "tile.hpp"

namespace mscx {

std::atomic<uint64_t> g_dmerges{0}, g_dfallbacks{0}, g_line_passes{0};

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr unsigned long long MAP_EMPTY = ~0ull;
constexpr uint8_t F_CANC = CPF_CANCELLED, F_BC = CPF_BCANC;
// device error bits (a set bit sends the grid to the host path, which
// reproduces the reference's behaviour, error messages included)
enum : uint32_t {
  DE_MAPDUP = 1,    // a cell inserted twice ("this cp is present in msc")
  DE_PAIRMISS = 2,  // a pair lookup the reference answers with a zero entry
  DE_LINEMISS = 4,  // a shared-line cell missing ("missing common bndry cp")
  DE_LONG = 8,      // an arc list longer than the workgroup sort takes
  DE_DIM = 16,      // an arc between cells of non-adjacent dimension
  DE_LOGIC = 32,    // an invariant the reference checks with logic_error
  DE_MUTUAL = 64,   // a non-mutual pairing met by merge_down
};
constexpr int SORT_SHORT = 64;     // lists sorted by one thread
constexpr int SORT_LDS = 8192;     // longest list a workgroup sorts in LDS
constexpr uint32_t LONGQ_CAP = 1u << 20;

struct DView {
  uint32_t *cell, *pair;
  float *fn;
  uint8_t *flags;
  uint32_t *off, *cnt, *arcs;
  unsigned long long *map;
  uint32_t mmask, n;
};

DView view(const DComplex &d) {
  return DView{d.cell, d.pair, d.fn, d.flags, d.off, d.cnt, d.arcs, d.map, d.mmask, d.n};
}

struct RectD {
  int x0, y0, x1, y1;
};
RectD rd(const Rect &r) { return RectD{r.x0, r.y0, r.x1, r.y1}; }

__device__ __forceinline__ int dcx(uint32_t c) { return (int16_t)(c & 0xFFFFu); }
__device__ __forceinline__ int dcy(uint32_t c) { return (int16_t)(c >> 16); }
__device__ __forceinline__ uint32_t ddim(uint32_t c) { return (c & 1u) + ((c >> 16) & 1u); }
__device__ __forceinline__ bool rcontains(const RectD &r, uint32_t c) {
  const int x = dcx(c), y = dcy(c);
  return x >= r.x0 && y >= r.y0 && x <= r.x1 && y <= r.y1;
}

__device__ __forceinline__ uint32_t dhash(uint32_t k) {
  k ^= k >> 16;
  k *= 0x7FEB352Du;
  k ^= k >> 15;
  k *= 0x846CA68Bu;
  k ^= k >> 16;
  return k;
}

// A slot is {key, value} as two little-endian words (the host CellMap's
// layout, so that a host copy of a complex takes the table as it is).
__device__ uint32_t map_find(const unsigned long long *map, uint32_t mask, uint32_t key) {
  uint32_t i = dhash(key) & mask;
  for (;;) {
    const unsigned long long e = map[i];
    if (e == MAP_EMPTY) return NONE;
    if ((uint32_t)e == key) return (uint32_t)(e >> 32);
    i = (i + 1) & mask;
  }
}

// false when key is already present
__device__ bool map_insert(unsigned long long *map, uint32_t mask, uint32_t key, uint32_t val) {
  uint32_t i = dhash(key) & mask;
  const unsigned long long v = ((unsigned long long)val << 32) | key;
  for (;;) {
    const unsigned long long old = atomicCAS(&map[i], MAP_EMPTY, v);
    if (old == MAP_EMPTY) return true;
    if ((uint32_t)old == key) return false;
    i = (i + 1) & mask;
  }
}

__device__ __forceinline__ void set_err(uint32_t *err, uint32_t bit) { atomicOr(err, bit); }

// ------------------------------------------------------------- list sorts

__global__ __launch_bounds__(256) void k_sort_short(const uint32_t *__restrict__ off, const uint32_t *__restrict__ cnt,
                                                    uint32_t *arcs, uint64_t nlists, uint32_t *longq,
                                                    uint32_t *nlong, uint32_t *err) {
  const uint64_t l = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (l >= nlists) return;
  const uint32_t c = cnt[l];
  if (c <= 1) return;
  if (c > SORT_SHORT) {
    const uint32_t q = atomicAdd(nlong, 1u);
    if (q < LONGQ_CAP) longq[q] = (uint32_t)l;
    else set_err(err, DE_LONG);
    return;
  }
  uint32_t *p = arcs + off[l];
  for (uint32_t i = 1; i < c; ++i) {
    const uint32_t v = p[i];
    uint32_t j = i;
    while (j > 0 && p[j - 1] > v) {
      p[j] = p[j - 1];
      --j;
    }
    p[j] = v;
  }
}

// one workgroup per long list (bitonic sort in LDS); lists longer than LDS
// holds go to a second queue that the host sorts
__global__ __launch_bounds__(256) void k_sort_long(const uint32_t *__restrict__ off, const uint32_t *__restrict__ cnt,
                                                   uint32_t *arcs, const uint32_t *__restrict__ longq,
                                                   const uint32_t *__restrict__ nlong, uint32_t *xlq, uint32_t *nxl,
                                                   uint32_t *err) {
  __shared__ uint32_t s[SORT_LDS];
  const uint32_t nq = min(*nlong, LONGQ_CAP);
  for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
    const uint32_t l = longq[q];
    const uint32_t c = cnt[l];
    if (c > SORT_LDS) {
      if (threadIdx.x == 0) {
        const uint32_t x = atomicAdd(nxl, 1u);
        if (x < LONGQ_CAP) xlq[x] = l;
        else set_err(err, DE_LONG);
      }
      continue;
    }
    uint32_t N = 1;
    while (N < c) N <<= 1;
    uint32_t *p = arcs + off[l];
    for (uint32_t i = threadIdx.x; i < N; i += 256) s[i] = i < c ? p[i] : NONE;
    __syncthreads();
    for (uint32_t k = 2; k <= N; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < N; i += 256) {
          const uint32_t ixj = i ^ j;
          if (ixj > i) {
            const uint32_t a = s[i], b = s[ixj];
            const bool up = (i & k) == 0;
            if ((a > b) == up) {
              s[i] = b;
              s[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t i = threadIdx.x; i < c; i += 256) p[i] = s[i];
    __syncthreads();
  }
}

// ------------------------------------------------------------- maps, misc

__global__ __launch_bounds__(256) void k_map_build(DView d, uint32_t from, uint32_t to, uint32_t *err) {
  const uint32_t i = from + blockIdx.x * 256 + threadIdx.x;
  if (i >= to) return;
  if (!map_insert(d.map, d.mmask, d.cell[i], i)) set_err(err, DE_MAPDUP);
}

// ------------------------------------------------------------- leaves

// writeout_connectivity (grid_dataset.cpp:1858-1907) from the critical
// block: nodes in raster order; a saddle's des = owners of its facets, asc =
// owners of its cofacets (both sorted); extrema count their saddles.
__global__ __launch_bounds__(256) void k_leaf_nodes(const uint32_t *__restrict__ crit, const float *__restrict__ fn,
                                                    const uint32_t *__restrict__ conn, const uint32_t *__restrict__ pr,
                                                    uint32_t K, DView o, uint32_t *err) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= K) return;
  const uint32_t c = crit[i];
  o.cell[i] = c;
  o.fn[i] = fn[i];
  const uint32_t p = pr[i];
  o.pair[i] = p;
  o.flags[i] = p != NONE ? F_BC : 0;
  if (ddim(c) != 1) return;
  const uint32_t *q = conn + 4 * (size_t)i;
  uint32_t nd = 0, na = 0;
  for (int k = 0; k < 2; ++k) {
    const uint32_t v = q[k];
    if (v == NONE) continue;
    ++nd;
    if (v >= K || ddim(crit[v]) != 0) {
      set_err(err, DE_DIM);
      continue;
    }
    atomicAdd(&o.cnt[2 * (size_t)v + 0], 1u);
  }
  for (int k = 2; k < 4; ++k) {
    const uint32_t v = q[k];
    if (v == NONE) continue;
    ++na;
    if (v >= K || ddim(crit[v]) != 2) {
      set_err(err, DE_DIM);
      continue;
    }
    atomicAdd(&o.cnt[2 * (size_t)v + 1], 1u);
  }
  o.cnt[2 * (size_t)i + 0] = na;
  o.cnt[2 * (size_t)i + 1] = nd;
}

__global__ __launch_bounds__(256) void k_leaf_arcs(const uint32_t *__restrict__ crit, const uint32_t *__restrict__ conn,
                                                   uint32_t K, DView o, uint32_t *cur) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= K || ddim(crit[i]) != 1) return;
  const uint32_t *q = conn + 4 * (size_t)i;
  for (int side = 0; side < 2; ++side) {  // 0: des (facets, minima), 1: asc (cofacets, maxima)
    uint32_t a = q[2 * side], b = q[2 * side + 1];
    if (a != NONE && b != NONE && b < a) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const int lk = side == 0 ? 1 : 0;  // the saddle's own list
    uint32_t *dst = o.arcs + o.off[2 * (size_t)i + lk];
    int w = 0;
    if (a != NONE) dst[w++] = a;
    if (b != NONE) dst[w++] = b;
    const uint32_t t2[2] = {a, b};
    for (int k = 0; k < 2; ++k) {
      const uint32_t v = t2[k];
      if (v == NONE || v >= K) continue;
      const int vk = side == 0 ? 0 : 1;  // minimum: asc list, maximum: des list
      const uint32_t pos = atomicAdd(&cur[2 * (size_t)v + vk], 1u);
      o.arcs[o.off[2 * (size_t)v + vk] + pos] = i;
    }
  }
}

// ------------------------------------------------------------- merge_up

__global__ __launch_bounds__(256) void k_mu_keep_a(DView a, uint32_t *f0) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j < a.n) f0[j] = (a.flags[j] & F_CANC) ? 0u : 1u;
}

// msc2's node j: the union index of msc1's non-cancelled node on the same
// cell (sh), or a new slot when it is itself not cancelled (newf)
__global__ __launch_bounds__(256) void k_mu_keep_b(DView a, DView b, const uint32_t *__restrict__ ex0, uint32_t *sh,
                                                   uint32_t *newf) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= b.n) return;
  const uint32_t k = map_find(a.map, a.mmask, b.cell[j]);
  if (k != NONE && !(a.flags[k] & F_CANC)) {
    sh[j] = ex0[k];
    newf[j] = 0;
  } else {
    sh[j] = NONE;
    newf[j] = (b.flags[j] & F_CANC) ? 0u : 1u;
  }
}

__global__ __launch_bounds__(256) void k_mu_nodes(DView a, DView b, const uint32_t *__restrict__ ex0, uint32_t *sh_at1,
                                                  const uint32_t *__restrict__ exb, uint32_t n1, DView o,
                                                  uint32_t *src0, uint32_t *src1) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t < a.n) {
    const uint32_t j = t;
    if (!(a.flags[j] & F_CANC)) {
      const uint32_t d = ex0[j];
      o.cell[d] = a.cell[j];
      o.fn[d] = a.fn[j];
      o.flags[d] = a.flags[j];
      o.pair[d] = NONE;
      src0[d] = j;
    }
  } else if (t - a.n < b.n) {
    const uint32_t j = t - a.n;
    const uint32_t s = sh_at1[j];
    uint32_t d = s;
    if (s == NONE) d = (b.flags[j] & F_CANC) ? NONE : n1 + exb[j];
    sh_at1[j] = d;  // becomes at1: msc2's index map into the union
    if (d == NONE) return;
    src1[d] = j;
    if (d >= n1) {
      o.cell[d] = b.cell[j];
      o.fn[d] = b.fn[j];
      o.flags[d] = b.flags[j];
      o.pair[d] = NONE;
    }
  }
}

// Arc copies of merge_up (grid_mscomplex.cpp:77-119): msc1's and msc2's
// lists of the union node's sources, cancelled targets dropped, msc2's
// shared-line to shared-line arcs dropped.  write = false counts.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_mu_arcs(DView a, DView b, const uint32_t *__restrict__ ex0,
                                                 const uint32_t *__restrict__ at1, const uint32_t *__restrict__ src0,
                                                 const uint32_t *__restrict__ src1, DView o, RectD ixn) {
  const uint32_t d = blockIdx.x * 256 + threadIdx.x;
  if (d >= o.n) return;
  const uint32_t j0 = src0[d], j1 = src1[d];
  const bool use1 = j1 != NONE && !(b.flags[j1] & F_CANC);
  const bool cmn = use1 && rcontains(ixn, b.cell[j1]);
  for (int k = 0; k < 2; ++k) {
    uint32_t c = 0;
    uint32_t *dst = WRITE ? o.arcs + o.off[2 * (size_t)d + k] : nullptr;
    if (j0 != NONE) {
      const uint32_t *p = a.arcs + a.off[2 * (size_t)j0 + k];
      const uint32_t e = a.cnt[2 * (size_t)j0 + k];
      for (uint32_t q = 0; q < e; ++q) {
        const uint32_t v = p[q];
        if (a.flags[v] & F_CANC) continue;
        if (WRITE) dst[c] = ex0[v];
        ++c;
      }
    }
    if (use1) {
      const uint32_t *p = b.arcs + b.off[2 * (size_t)j1 + k];
      const uint32_t e = b.cnt[2 * (size_t)j1 + k];
      for (uint32_t q = 0; q < e; ++q) {
        const uint32_t v = p[q];
        if (cmn && rcontains(ixn, b.cell[v])) continue;
        if (b.flags[v] & F_CANC) continue;
        if (WRITE) dst[c] = at1[v];
        ++c;
      }
    }
    if (!WRITE) o.cnt[2 * (size_t)d + k] = c;
  }
}

// pair indices (grid_mscomplex.cpp:90-95): msc1's lookup, then msc2's,
// which wins for a shared node; a lookup that misses would insert a zero
// entry in the reference's map
__global__ __launch_bounds__(256) void k_mu_pairs(DView a, DView b, const uint32_t *__restrict__ src0,
                                                  const uint32_t *__restrict__ src1, DView o, uint32_t *err) {
  const uint32_t d = blockIdx.x * 256 + threadIdx.x;
  if (d >= o.n) return;
  uint32_t val = NONE;
  const uint32_t j0 = src0[d], j1 = src1[d];
  if (j0 != NONE && (a.flags[j0] & F_BC)) {
    const uint32_t p = a.pair[j0];
    const uint32_t r = p < a.n ? map_find(o.map, o.mmask, a.cell[p]) : NONE;
    if (r == NONE) set_err(err, DE_PAIRMISS);
    val = r;
  }
  if (j1 != NONE && !(b.flags[j1] & F_CANC) && (b.flags[j1] & F_BC)) {
    const uint32_t p = b.pair[j1];
    const uint32_t r = p < b.n ? map_find(o.map, o.mmask, b.cell[p]) : NONE;
    if (r == NONE) set_err(err, DE_PAIRMISS);
    val = r;
  }
  o.pair[d] = val;
}

// ------------------------------------------------------------- line passes

__device__ __forceinline__ void bm_set(uint32_t *bm, uint32_t i) { atomicOr(&bm[i >> 5], 1u << (i & 31)); }

__device__ void mark_lists(const DView &d, uint32_t x, uint32_t *bm, bool pairs_of_bc) {
  for (int k = 0; k < 2; ++k) {
    const uint32_t *p = d.arcs + d.off[2 * (size_t)x + k];
    const uint32_t e = d.cnt[2 * (size_t)x + k];
    for (uint32_t q = 0; q < e; ++q) {
      const uint32_t v = p[q];
      bm_set(bm, v);
      if (pairs_of_bc && (d.flags[v] & F_BC)) {
        const uint32_t pv = d.pair[v];
        if (pv < d.n) bm_set(bm, pv);
      }
    }
  }
}

// The nodes whose lists a line pass reads or writes: every line cell's node;
// for merge_up the candidates' pairs and the neighbours of both; for
// merge_down the cancelled line nodes' pairs, the neighbours of both, and
// the pairs of the boundary-cancellable neighbours (uncancel_pairs expands
// through them).
__global__ __launch_bounds__(256) void k_line_mark(DView d, RectD ixn, int up, uint32_t *bm, uint32_t *err) {
  const int w = ixn.x1 - ixn.x0 + 1;
  const int nl = w * (ixn.y1 - ixn.y0 + 1);
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nl) return;
  const int x = ixn.x0 + t % w, y = ixn.y0 + t / w;
  const uint32_t c = (uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16);
  const uint32_t si = map_find(d.map, d.mmask, c);
  if (si == NONE) {
    set_err(err, DE_LINEMISS);
    return;
  }
  bm_set(bm, si);
  const uint8_t f = d.flags[si];
  const bool cand = up ? ((f & F_BC) && !(f & F_CANC)) : (f & F_CANC) != 0;
  if (!cand) return;
  const uint32_t pi = d.pair[si];
  if (pi >= d.n) {
    set_err(err, DE_LOGIC);
    return;
  }
  bm_set(bm, pi);
  mark_lists(d, si, bm, !up);
  mark_lists(d, pi, bm, !up);
}

__global__ __launch_bounds__(256) void k_mark_meta(DView d, const uint32_t *__restrict__ ids, const uint32_t *nids,
                                                   uint32_t *bm) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= *nids) return;
  const uint32_t x = ids[t];
  bm_set(bm, x);
  mark_lists(d, x, bm, false);
}

__global__ __launch_bounds__(256) void k_bm_count(const uint32_t *__restrict__ bm, uint32_t nw, uint32_t *cnt) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < nw) cnt[i] = __popc(bm[i]);
}

__global__ __launch_bounds__(256) void k_bm_write(const uint32_t *__restrict__ bm, uint32_t nw,
                                                  const uint32_t *__restrict__ off, uint32_t *ids) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nw) return;
  uint32_t w = bm[i], o = off[i];
  while (w) {
    const int b = __ffs(w) - 1;
    ids[o++] = i * 32 + b;
    w &= w - 1;
  }
}

// meta records: cell, flags, pair (3 words per id)
__global__ __launch_bounds__(256) void k_pack_meta(DView d, const uint32_t *__restrict__ ids, uint32_t n,
                                                   uint32_t *out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const uint32_t x = ids[t];
  out[3 * (size_t)t] = d.cell[x];
  out[3 * (size_t)t + 1] = d.flags[x];
  out[3 * (size_t)t + 2] = d.pair[x];
}

// list records of ids: na, nd, asc..., des... (sizes first, then writes)
template <bool WRITE>
__global__ __launch_bounds__(256) void k_pack_lists(DView d, const uint32_t *__restrict__ ids, uint32_t n,
                                                    uint32_t *sz, const uint32_t *__restrict__ rec_off, uint32_t *out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const uint32_t x = ids ? ids[t] : t;
  const uint32_t na = d.cnt[2 * (size_t)x], nd = d.cnt[2 * (size_t)x + 1];
  if (!WRITE) {
    sz[t] = 2 + na + nd;
    return;
  }
  uint32_t *w = out + rec_off[t];
  w[0] = na;
  w[1] = nd;
  const uint32_t *pa = d.arcs + d.off[2 * (size_t)x], *pd = d.arcs + d.off[2 * (size_t)x + 1];
  for (uint32_t q = 0; q < na; ++q) w[2 + q] = pa[q];
  for (uint32_t q = 0; q < nd; ++q) w[2 + na + q] = pd[q];
}

// patch records written back by a line pass: node, flags, na, nd, offset
// of its new lists in the pool (the lists follow in the patch words)
__global__ __launch_bounds__(256) void k_apply_patch(DView d, const uint32_t *__restrict__ words,
                                                     const uint32_t *__restrict__ rec_off, uint32_t nrec) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nrec) return;
  const uint32_t *r = words + rec_off[t];
  const uint32_t x = r[0], na = r[2], nd = r[3], base = r[4];
  d.flags[x] = (uint8_t)r[1];
  d.off[2 * (size_t)x] = base;
  d.cnt[2 * (size_t)x] = na;
  d.off[2 * (size_t)x + 1] = base + na;
  d.cnt[2 * (size_t)x + 1] = nd;
  for (uint32_t q = 0; q < na + nd; ++q) d.arcs[base + q] = r[5 + q];
}

// ------------------------------------------------------------- merge_down

// at[j]: the child's index of P's cell j
__global__ __launch_bounds__(256) void k_md_lookup(DView P, DView m, uint32_t *at, uint8_t *in0,
                                                   unsigned long long *key, uint32_t *err) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= P.n) return;
  if (P.flags[j] & F_CANC) set_err(err, DE_LOGIC);  // "all cps ought to be uncancelled by now"
  const uint32_t r = map_find(m.map, m.mmask, P.cell[j]);
  at[j] = r;
  in0[j] = r != NONE;
  key[j] = ~0ull;
}

// A boundary-cancellable node of P whose pair does not point back (the
// simplification re-pairs the nodes it cancels, leaving the old partner's
// pair field behind) makes the reference's walk order-dependent: such a
// node, its pair and its pair's partner form a component that the host
// walks in index order.  Longer pair chains are not followed (fallback).
__global__ __launch_bounds__(256) void k_md_cmark(DView P, uint32_t *bm, uint32_t *err) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= P.n || !(P.flags[j] & F_BC)) return;
  const uint32_t sp = P.pair[j];
  if (sp >= P.n || sp == j) {
    set_err(err, DE_MUTUAL);
    return;
  }
  if (P.pair[sp] == j) return;
  bm_set(bm, j);
  bm_set(bm, sp);
  if (!(P.flags[sp] & F_BC)) return;
  const uint32_t s2 = P.pair[sp];
  if (s2 >= P.n || P.pair[s2] != sp) {
    set_err(err, DE_MUTUAL);
    return;
  }
  bm_set(bm, s2);
}

// complex-component records for the host walk: P flags, P pair, present in
// the child, the child's boundary-cancellable bit
__global__ __launch_bounds__(256) void k_md_cpack(DView P, DView m, const uint32_t *__restrict__ ids, uint32_t nc,
                                                  const uint32_t *__restrict__ at, uint32_t *out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  const uint32_t x = ids[t], a = at[x];
  out[4 * (size_t)t] = P.flags[x];
  out[4 * (size_t)t + 1] = P.pair[x];
  out[4 * (size_t)t + 2] = a != NONE;
  out[4 * (size_t)t + 3] = a != NONE ? (m.flags[a] & F_BC) : 0u;
}

// processed steps of the reference's walk (grid_mscomplex.cpp:207-224): a
// boundary-cancellable node whose cell or whose pair's cell is in the child;
// complex components take the host walk's answer (res: processed bit 0,
// boundary-cancellable set bit 1, pair written: P index in the next word)
__global__ __launch_bounds__(256) void k_md_proc(DView P, const uint8_t *__restrict__ in0, uint8_t *proc,
                                                 uint8_t *cplx) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= P.n) return;
  cplx[j] = 0;
  uint8_t r = 0;
  if (P.flags[j] & F_BC) {
    const uint32_t sp = P.pair[j];
    r = sp < P.n && (in0[j] || in0[sp]);
  }
  proc[j] = r;
}

__global__ __launch_bounds__(256) void k_md_cset(const uint32_t *__restrict__ ids, uint32_t nc,
                                                 const uint32_t *__restrict__ res, uint8_t *proc, uint8_t *cplx) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  const uint32_t x = ids[t];
  cplx[x] = 1;
  proc[x] = res[2 * (size_t)t] & 1u;
}

// First reference of every node merge_down replicates into the child
// (grid_mscomplex.cpp:207-268 walks P's boundary-cancellable nodes in
// order and replicates the node, its pair and its arcs' ends on first
// sight): key = (j << 32) | position, position 0 the node itself, 1 its
// pair, 2 + q the q-th entry of asc ++ des.
__global__ __launch_bounds__(256) void k_md_keys(DView P, const uint8_t *__restrict__ in0,
                                                 const uint8_t *__restrict__ proc, unsigned long long *key,
                                                 uint32_t *err) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= P.n || !proc[j]) return;
  const uint32_t sp = P.pair[j];
  const unsigned long long kj = (unsigned long long)j << 32;
  if (!in0[j]) atomicMin(&key[j], kj);
  if (!in0[sp]) atomicMin(&key[sp], kj | 1ull);
  uint32_t pos = 2;
  for (int k = 0; k < 2; ++k) {
    const uint32_t *p = P.arcs + P.off[2 * (size_t)j + k];
    const uint32_t e = P.cnt[2 * (size_t)j + k];
    for (uint32_t q = 0; q < e; ++q, ++pos) {
      const uint32_t v = p[q];
      if (P.flags[v] & F_BC) set_err(err, DE_LOGIC);  // "only non cancellable cps must be remaining"
      if (!in0[v]) atomicMin(&key[v], kj | pos);
    }
  }
}

// replicas first referenced by j, counted (ASSIGN = false) or numbered
template <bool ASSIGN>
__global__ __launch_bounds__(256) void k_md_reps(DView P, const uint8_t *__restrict__ in0,
                                                 const uint8_t *__restrict__ proc,
                                                 const unsigned long long *__restrict__ key, uint32_t *cntj,
                                                 const uint32_t *__restrict__ base, uint32_t n0, uint32_t *at,
                                                 uint32_t *rep_src) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= P.n) return;
  uint32_t c = 0;
  if (proc[j]) {
    const unsigned long long kj = (unsigned long long)j << 32;
    const uint32_t b0 = ASSIGN ? base[j] : 0;
    auto take = [&](uint32_t v, unsigned long long kk) {
      if (!in0[v] && key[v] == kk) {
        if (ASSIGN) {
          at[v] = n0 + b0 + c;
          rep_src[b0 + c] = v;
        }
        ++c;
      }
    };
    take(j, kj);
    take(P.pair[j], kj | 1ull);
    uint32_t pos = 2;
    for (int k = 0; k < 2; ++k) {
      const uint32_t *p = P.arcs + P.off[2 * (size_t)j + k];
      const uint32_t e = P.cnt[2 * (size_t)j + k];
      for (uint32_t q = 0; q < e; ++q, ++pos) take(p[q], kj | pos);
    }
  }
  if (!ASSIGN) cntj[j] = c;
}

__global__ __launch_bounds__(256) void k_md_rep_fill(DView P, DView m, const uint32_t *__restrict__ rep_src, uint32_t R,
                                                     uint32_t n0, uint32_t *err) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  const uint32_t v = rep_src[r], i = n0 + r;
  m.cell[i] = P.cell[v];
  m.fn[i] = P.fn[v];
  m.flags[i] = P.flags[v];
  m.pair[i] = NONE;
  if (!map_insert(m.map, m.mmask, P.cell[v], i)) set_err(err, DE_MAPDUP);
}

// pair fields (grid_mscomplex.cpp:230-248) of each processed mutual pair
// j < sp: the reference's two steps (j, then sp) reduce to one test
__global__ __launch_bounds__(256) void k_md_pairs(DView P, DView m, const uint8_t *__restrict__ in0,
                                                  const uint8_t *__restrict__ proc, const uint8_t *__restrict__ cplx,
                                                  const uint32_t *__restrict__ at) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= P.n || !proc[j] || cplx[j]) return;
  const uint32_t sp = P.pair[j];
  if (sp < j) return;
  const uint32_t dj = at[j], ds = at[sp];
  if (dj == NONE || ds == NONE) return;  // both are present or replicated by now
  const bool set = !in0[j] || !in0[sp] || !(m.flags[dj] & F_BC) || !(m.flags[ds] & F_BC);
  if (!set) return;
  m.flags[dj] |= F_BC;
  m.flags[ds] |= F_BC;
  m.pair[dj] = ds;
  m.pair[ds] = dj;
}

// the host walk's pair fields for complex components
__global__ __launch_bounds__(256) void k_md_pairs_cplx(DView m, const uint32_t *__restrict__ ids, uint32_t nc,
                                                       const uint32_t *__restrict__ res,
                                                       const uint32_t *__restrict__ at) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  const uint32_t d = at[ids[t]];
  if (d == NONE) return;
  const uint32_t r = res[2 * (size_t)t], pv = res[2 * (size_t)t + 1];
  if (r & 2u) m.flags[d] |= F_BC;
  if (pv != NONE) m.pair[d] = at[pv];
}

__global__ __launch_bounds__(256) void k_md_inv(uint32_t nP, const uint32_t *__restrict__ at, uint32_t *inv) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j < nP && at[j] != NONE) inv[at[j]] = j;
}

// The child's lists after merge_down (grid_mscomplex.cpp:250-318): a node
// present in P takes P's lists translated into the child (ends absent from
// the child dropped); any other node keeps its own.  WRITE = false counts.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_md_arcs(DView P, DView m, const uint32_t *__restrict__ at,
                                                 const uint32_t *__restrict__ inv, uint32_t *ncnt,
                                                 const uint32_t *__restrict__ noff, uint32_t *narcs) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m.n) return;
  const uint32_t j = inv[i];
  for (int k = 0; k < 2; ++k) {
    uint32_t *dst = WRITE ? narcs + noff[2 * (size_t)i + k] : nullptr;
    uint32_t c = 0;
    if (j != NONE) {
      const uint32_t *p = P.arcs + P.off[2 * (size_t)j + k];
      const uint32_t e = P.cnt[2 * (size_t)j + k];
      for (uint32_t q = 0; q < e; ++q) {
        const uint32_t mv = at[p[q]];
        if (mv == NONE) continue;
        if (WRITE) dst[c] = mv;
        ++c;
      }
    } else {
      const uint32_t *p = m.arcs + m.off[2 * (size_t)i + k];
      const uint32_t e = m.cnt[2 * (size_t)i + k];
      if (WRITE)
        for (uint32_t q = 0; q < e; ++q) dst[q] = p[q];
      c = e;
    }
    if (!WRITE) ncnt[2 * (size_t)i + k] = c;
  }
}

// ------------------------------------------------------------- host copies

template <bool WRITE>
__global__ __launch_bounds__(256) void k_pack_all(DView d, uint32_t *sz, const uint32_t *__restrict__ rec_off,
                                                  uint32_t *out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= d.n) return;
  const uint32_t na = d.cnt[2 * (size_t)i], nd = d.cnt[2 * (size_t)i + 1];
  if (!WRITE) {
    sz[i] = 6 + na + nd;
    return;
  }
  uint32_t *w = out + rec_off[i];
  w[0] = d.cell[i];
  w[1] = __float_as_uint(d.fn[i]);
  w[2] = d.pair[i];
  w[3] = d.flags[i];
  w[4] = na;
  w[5] = nd;
  const uint32_t *pa = d.arcs + d.off[2 * (size_t)i], *pd = d.arcs + d.off[2 * (size_t)i + 1];
  for (uint32_t q = 0; q < na; ++q) w[6 + q] = pa[q];
  for (uint32_t q = 0; q < nd; ++q) w[6 + na + q] = pd[q];
}

__global__ __launch_bounds__(256) void k_flags_from_u32(const uint32_t *__restrict__ in, uint8_t *out, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = (uint8_t)in[i];
}

// ------------------------------------------------------------- host side

// at least one block: kernels bound-check their index
inline unsigned blocks(uint64_t n) { return n ? (unsigned)((n + 255) / 256) : 1u; }

// Temporary device buffers of one operation: returned to the pool only after
// the stream has drained (the destructor synchronises), so that another
// thread's allocation never receives a block a queued kernel still uses.
struct Temps {
  hipStream_t s;
  std::vector<void *> v;
  explicit Temps(hipStream_t st) : s(st) {}
  template <class T>
  T *get(uint64_t n) {
    void *p = device_pool().alloc(std::max<uint64_t>(n, 1) * sizeof(T));
    v.push_back(p);
    return (T *)p;
  }
  ~Temps() {
    (void)hipStreamSynchronize(s);
    for (void *p : v) device_pool().free(p);
  }
};

// exclusive scan of n words; the total lands in *d_total when given
void dscan(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *d_total, hipStream_t s, Temps &T) {
  dev_scan(in, out, n, d_total, T.get<uint32_t>(scan_temp_words(n)), s);
}

void sort_lists(DComplex &d, hipStream_t s, Temps &T, uint32_t *d_err) {
  const uint64_t nl = 2 * (uint64_t)d.n;
  if (!nl) return;
  uint32_t *longq = T.get<uint32_t>(LONGQ_CAP), *xlq = T.get<uint32_t>(LONGQ_CAP);
  uint32_t *nlong = T.get<uint32_t>(2);
  MSCX_HIP(hipMemsetAsync(nlong, 0, 8, s));
  k_sort_short<<<blocks(nl), 256, 0, s>>>(d.off, d.cnt, d.arcs, nl, longq, nlong, d_err);
  k_sort_long<<<256, 256, 0, s>>>(d.off, d.cnt, d.arcs, longq, nlong, xlq, nlong + 1, d_err);
  uint32_t hx = 0;
  MSCX_HIP(hipMemcpyAsync(&hx, nlong + 1, 4, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipStreamSynchronize(s));
  hx = std::min(hx, LONGQ_CAP);
  if (!hx) return;
  // lists longer than a workgroup's LDS (dense strangulation multiplicities):
  // sorted on the host
  std::vector<uint32_t> ids(hx);
  MSCX_HIP(hipMemcpy(ids.data(), xlq, (size_t)hx * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> v;
  for (uint32_t l : ids) {
    uint32_t o = 0, c = 0;
    MSCX_HIP(hipMemcpy(&o, d.off + l, 4, hipMemcpyDeviceToHost));
    MSCX_HIP(hipMemcpy(&c, d.cnt + l, 4, hipMemcpyDeviceToHost));
    v.resize(c);
    MSCX_HIP(hipMemcpy(v.data(), d.arcs + o, (size_t)c * 4, hipMemcpyDeviceToHost));
    std::sort(v.begin(), v.end());
    MSCX_HIP(hipMemcpy(d.arcs + o, v.data(), (size_t)c * 4, hipMemcpyHostToDevice));
  }
}

uint32_t map_slots(uint32_t ncap) {
  uint32_t m = 16;
  while (m < 2ull * ncap) m <<= 1;
  return m;
}

// node arrays and map for ncap nodes (contents undefined, map empty)
void alloc_nodes(DComplex &d, uint32_t ncap, hipStream_t s) {
  const size_t a4 = align_up((size_t)ncap * 4, 256), a1 = align_up((size_t)ncap, 256);
  const size_t a8 = align_up((size_t)ncap * 8, 256);
  char *p = (char *)device_pool().alloc(3 * a4 + a1 + 2 * a8);
  d.node_mem = p;
  d.cell = (uint32_t *)p;
  d.pair = (uint32_t *)(p + a4);
  d.fn = (float *)(p + 2 * a4);
  d.flags = (uint8_t *)(p + 3 * a4);
  d.off = (uint32_t *)(p + 3 * a4 + a1);
  d.cnt = (uint32_t *)(p + 3 * a4 + a1 + a8);
  d.ncap = ncap;
  const uint32_t slots = map_slots(ncap);
  d.map_mem = device_pool().alloc((size_t)slots * 8);
  d.map = (unsigned long long *)d.map_mem;
  d.mmask = slots - 1;
  MSCX_HIP(hipMemsetAsync(d.map, 0xFF, (size_t)slots * 8, s));
}

void alloc_arcs(DComplex &d, uint64_t acap) {
  MSCX_REQUIRE(acap < (1ull << 32), MSCX_ENOMEM, "arc pool of a device complex exceeds 2^32 entries");
  d.arc_mem = device_pool().alloc(std::max<uint64_t>(acap, 1) * 4);
  d.arcs = (uint32_t *)d.arc_mem;
  d.acap = acap;
}

// room for `more` arc words past aused (existing lists are kept)
void reserve_arcs(DComplex &d, uint64_t more, hipStream_t s) {
  if (d.aused + more <= d.acap) return;
  const uint64_t want = d.aused + more + (d.aused + more) / 8 + 1024;
  void *old = d.arc_mem;
  uint32_t *oa = d.arcs;
  alloc_arcs(d, want);
  if (d.aused) MSCX_HIP(hipMemcpyAsync(d.arcs, oa, d.aused * 4, hipMemcpyDeviceToDevice, s));
  MSCX_HIP(hipStreamSynchronize(s));
  device_pool().free(old);
}

uint32_t read_u32(const uint32_t *dev, hipStream_t s) {
  uint32_t v = 0;
  MSCX_HIP(hipMemcpyAsync(&v, dev, 4, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipStreamSynchronize(s));
  return v;
}

void check_err(const uint32_t *d_err, hipStream_t s) {
  if (const uint32_t e = read_u32(d_err, s)) throw DFallback("device check bits " + std::to_string(e));
}

// pinned host staging for the copies of this module, leased per call so
// that concurrent copies (the leaves of one level) do not wait on each other
std::mutex g_pin_mu;
std::vector<Scratch *> g_pins;
struct PinLease {
  Scratch *sc;
  PinLease() {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    if (g_pins.empty()) {
      sc = new Scratch;
    } else {
      sc = g_pins.back();
      g_pins.pop_back();
    }
  }
  ~PinLease() {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pins.push_back(sc);
  }
  PinLease(const PinLease &) = delete;
  PinLease &operator=(const PinLease &) = delete;
};
// copies up to this size go through pinned staging; larger ones (a
// 16384^2 root) straight into pageable memory
constexpr size_t PIN_MAX = (size_t)256 << 20;

// MSCX_DEBUG_MERGE in the environment: per-call timings on stderr
bool dbg_on() {
  static const bool on = getenv("MSCX_DEBUG_MERGE") != nullptr;
  return on;
}
struct DbgClock {
  const char *what;
  uint32_t n;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now(), last = t0;
  std::string marks;
  DbgClock(const char *w, uint32_t nn) : what(w), n(nn) {}
  void mark(const char *m) {
    if (!dbg_on()) return;
    const auto t = std::chrono::steady_clock::now();
    char b[64];
    snprintf(b, sizeof b, " %s %.3f", m, std::chrono::duration<double, std::milli>(t - last).count());
    marks += b;
    last = t;
  }
  ~DbgClock() {
    if (!dbg_on()) return;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    fprintf(stderr, "[dmerge] %s n=%u %.3f ms:%s\n", what, n, ms, marks.c_str());
  }
};

// host complex of the nodes a line pass touches
struct LineSub {
  std::vector<uint32_t> meta;  // global ids, sorted (every node in the sub-complex)
  std::vector<uint32_t> full;  // global ids, sorted (nodes with lists)
  Complex c;
  std::vector<uint8_t> flags0;
  std::vector<ConnSet> asc0, des0;  // lists of full nodes before the pass
};

uint32_t local_of(const std::vector<uint32_t> &ids, uint32_t g) {
  auto it = std::lower_bound(ids.begin(), ids.end(), g);
  return (it != ids.end() && *it == g) ? (uint32_t)(it - ids.begin()) : NONE;
}

// Gathers the sub-complex a line pass needs, runs the pass on the host and
// writes the changed nodes back.
void line_pass(DComplex &d, const Rect &ixn, bool up, hipStream_t s) {
  g_line_passes++;
  DbgClock dc(up ? "line_up" : "line_down", d.n);
  Temps T(s);
  const uint32_t nw = (d.n + 31) / 32;
  uint32_t *bm = T.get<uint32_t>(nw), *bm2 = T.get<uint32_t>(nw);
  uint32_t *woff = T.get<uint32_t>(nw);
  uint32_t *ids_full = T.get<uint32_t>(d.n), *ids_meta = T.get<uint32_t>(d.n);
  uint32_t *dn = T.get<uint32_t>(4);  // err, nfull, nmeta
  MSCX_HIP(hipMemsetAsync(bm, 0, (size_t)nw * 4, s));
  MSCX_HIP(hipMemsetAsync(bm2, 0, (size_t)nw * 4, s));
  MSCX_HIP(hipMemsetAsync(dn, 0, 16, s));
  const int nl = (ixn.x1 - ixn.x0 + 1) * (ixn.y1 - ixn.y0 + 1);
  const DView v = view(d);
  k_line_mark<<<blocks(nl), 256, 0, s>>>(v, rd(ixn), up ? 1 : 0, bm, dn);
  k_bm_count<<<blocks(nw), 256, 0, s>>>(bm, nw, woff);
  dscan(woff, woff, nw, dn + 1, s, T);
  k_bm_write<<<blocks(nw), 256, 0, s>>>(bm, nw, woff, ids_full);
  k_mark_meta<<<blocks(d.n), 256, 0, s>>>(v, ids_full, dn + 1, bm2);
  k_bm_count<<<blocks(nw), 256, 0, s>>>(bm2, nw, woff);
  dscan(woff, woff, nw, dn + 2, s, T);
  k_bm_write<<<blocks(nw), 256, 0, s>>>(bm2, nw, woff, ids_meta);
  uint32_t hn[4];
  MSCX_HIP(hipMemcpyAsync(hn, dn, 16, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipStreamSynchronize(s));
  // (a missing line cell: the host path raises the reference's error)
  if (hn[0]) throw DFallback("line pass check bits " + std::to_string(hn[0]));
  const uint32_t nfull = hn[1], nmeta = hn[2];
  dc.mark("mark");
  uint32_t *meta_rec = T.get<uint32_t>(3 * (uint64_t)nmeta);
  uint32_t *rsz = T.get<uint32_t>(nfull), *roff = T.get<uint32_t>(nfull);
  uint32_t *dtot = T.get<uint32_t>(1);
  k_pack_meta<<<blocks(nmeta), 256, 0, s>>>(v, ids_meta, nmeta, meta_rec);
  k_pack_lists<false><<<blocks(nfull), 256, 0, s>>>(v, ids_full, nfull, rsz, nullptr, nullptr);
  dscan(rsz, roff, nfull, dtot, s, T);
  const uint32_t lwords = read_u32(dtot, s);
  uint32_t *lrec = T.get<uint32_t>(lwords);
  k_pack_lists<true><<<blocks(nfull), 256, 0, s>>>(v, ids_full, nfull, nullptr, roff, lrec);

  LineSub L;
  L.meta.resize(nmeta);
  L.full.resize(nfull);
  std::vector<uint32_t> mrec(3 * (size_t)nmeta), lw(lwords);
  MSCX_HIP(hipMemcpyAsync(L.meta.data(), ids_meta, (size_t)nmeta * 4, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipMemcpyAsync(L.full.data(), ids_full, (size_t)nfull * 4, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipMemcpyAsync(mrec.data(), meta_rec, mrec.size() * 4, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipMemcpyAsync(lw.data(), lrec, (size_t)lwords * 4, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipStreamSynchronize(s));

  dc.mark("gather");
  Complex &c = L.c;
  c.rect = d.rect;
  c.ext = d.ext;
  c.cell.resize(nmeta);
  c.fn.assign(nmeta, 0.0f);
  c.pair.resize(nmeta);
  c.flags.resize(nmeta);
  c.asc.resize(nmeta);
  c.des.resize(nmeta);
  c.idx.reserve(nmeta);
  for (uint32_t i = 0; i < nmeta; ++i) {
    c.cell[i] = mrec[3 * (size_t)i];
    c.flags[i] = (uint8_t)mrec[3 * (size_t)i + 1];
    const uint32_t gp = mrec[3 * (size_t)i + 2];
    c.pair[i] = gp == NONE ? NONE : local_of(L.meta, gp);
    c.idx.insert(c.cell[i], i);
  }
  std::vector<uint32_t> floc(nfull);
  size_t p = 0;
  for (uint32_t k = 0; k < nfull; ++k) {
    const uint32_t i = local_of(L.meta, L.full[k]);
    floc[k] = i;
    const uint32_t na = lw[p], nd = lw[p + 1];
    p += 2;
    for (uint32_t q = 0; q < na; ++q) c.asc[i].push_sorted(local_of(L.meta, lw[p + q]));
    p += na;
    for (uint32_t q = 0; q < nd; ++q) c.des[i].push_sorted(local_of(L.meta, lw[p + q]));
    p += nd;
  }
  L.flags0 = c.flags;
  L.asc0.resize(nfull);
  L.des0.resize(nfull);
  for (uint32_t k = 0; k < nfull; ++k) {
    L.asc0[k] = c.asc[floc[k]];
    L.des0[k] = c.des[floc[k]];
  }
  std::vector<uint8_t> isfull(nmeta, 0);
  for (uint32_t k = 0; k < nfull; ++k) isfull[floc[k]] = 1;

  dc.mark("build");
  if (up) merge_up_line(c, ixn);
  else merge_down_line(c, ixn);
  dc.mark("pass");

  // only the full nodes may change (their lists hold every node a pass
  // touches); anything else means the sub-complex was not closed
  for (uint32_t i = 0; i < nmeta; ++i)
    if (!isfull[i] && (c.flags[i] != L.flags0[i] || c.asc[i].size() || c.des[i].size()))
      throw DFallback("line pass touched a node outside its sub-complex");
  auto same = [](const ConnSet &a, const ConnSet &b) {
    return a.size() == b.size() && std::equal(a.begin(), a.end(), b.begin());
  };
  std::vector<uint32_t> pw, pro;  // patch words, record offsets
  uint64_t add = 0;
  for (uint32_t k = 0; k < nfull; ++k) {
    const uint32_t i = floc[k];
    if (c.flags[i] == L.flags0[i] && same(c.asc[i], L.asc0[k]) && same(c.des[i], L.des0[k])) continue;
    pro.push_back((uint32_t)pw.size());
    const uint32_t na = c.asc[i].size(), nd = c.des[i].size();
    pw.push_back(L.full[k]);
    pw.push_back(c.flags[i]);
    pw.push_back(na);
    pw.push_back(nd);
    pw.push_back((uint32_t)(d.aused + add));
    for (uint32_t x : c.asc[i]) pw.push_back(L.meta[x]);
    for (uint32_t x : c.des[i]) pw.push_back(L.meta[x]);
    add += na + nd;
  }
  if (pro.empty()) return;
  reserve_arcs(d, add, s);
  // reserve_arcs may have moved the pool: the records' offsets are pool
  // offsets, unchanged by the move
  uint32_t *dw = T.get<uint32_t>(pw.size()), *dro = T.get<uint32_t>(pro.size());
  MSCX_HIP(hipMemcpyAsync(dw, pw.data(), pw.size() * 4, hipMemcpyHostToDevice, s));
  MSCX_HIP(hipMemcpyAsync(dro, pro.data(), pro.size() * 4, hipMemcpyHostToDevice, s));
  k_apply_patch<<<blocks(pro.size()), 256, 0, s>>>(view(d), dw, dro, (uint32_t)pro.size());
  d.aused += add;
  MSCX_HIP(hipStreamSynchronize(s));  // pw / pro are host vectors
}

}  // namespace

// ------------------------------------------------------------------ API

namespace {
std::mutex g_stream_mu;
std::vector<hipStream_t> g_streams;  // idle pooled streams
}  // namespace

StreamLease::StreamLease() {
  {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if (!g_streams.empty()) {
      s = g_streams.back();
      g_streams.pop_back();
      return;
    }
  }
  MSCX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
}

StreamLease::~StreamLease() {
  if (!s) return;
  (void)hipStreamSynchronize(s);
  std::lock_guard<std::mutex> lk(g_stream_mu);
  g_streams.push_back(s);
}

DComplex::~DComplex() {
  device_pool().free(node_mem);
  device_pool().free(arc_mem);
  device_pool().free(map_mem);
}

size_t DComplex::device_bytes() const {
  return (size_t)ncap * (3 * 4 + 1 + 16) + (size_t)acap * 4 + ((size_t)mmask + 1) * 8;
}

void dcomplex_from_tile(const mscx_tile *t, const Rect &rect, const Rect &ext, DComplex &o, hipStream_t s) {
  const uint32_t K = t->K;
  const char *b = (const char *)t->crit_block.get() + t->crit_off;
  const CritLayout L(K);
  const uint32_t *crit = (const uint32_t *)b, *conn = (const uint32_t *)(b + L.conn);
  const uint32_t *pr = (const uint32_t *)(b + L.pair);
  const float *fn = (const float *)(b + L.fn);
  o.rect = rect;
  o.ext = ext;
  alloc_nodes(o, (uint32_t)Complex::with_headroom(K), s);
  o.n = K;
  Temps T(s);
  uint32_t *err = T.get<uint32_t>(2);
  MSCX_HIP(hipMemsetAsync(err, 0, 8, s));
  MSCX_HIP(hipMemsetAsync(o.cnt, 0, (size_t)2 * K * 4, s));
  const DView v = view(o);
  k_leaf_nodes<<<blocks(K), 256, 0, s>>>(crit, fn, conn, pr, K, v, err);
  dscan(o.cnt, o.off, 2 * (uint64_t)K, err + 1, s, T);
  k_map_build<<<blocks(K), 256, 0, s>>>(v, 0, K, err);
  const uint32_t words = read_u32(err + 1, s);
  alloc_arcs(o, words + words / 8 + 256);
  o.aused = words;
  uint32_t *cur = T.get<uint32_t>(2 * (uint64_t)K);
  MSCX_HIP(hipMemsetAsync(cur, 0, (size_t)2 * K * 4, s));
  k_leaf_arcs<<<blocks(K), 256, 0, s>>>(crit, conn, K, view(o), cur);
  sort_lists(o, s, T, err);
  check_err(err, s);
}

void dmerge_up(const DComplex &A, const DComplex &B, DComplex &O, hipStream_t s) {
  DbgClock dc("merge_up", A.n + B.n);
  Rect ixn;
  merge_ixn(A.rect, B.rect, ixn);
  merge_rects(A.rect, A.ext, B.rect, B.ext, O.rect, O.ext);
  Temps T(s);
  uint32_t *dn = T.get<uint32_t>(4);  // err, n1, nb, arcs
  MSCX_HIP(hipMemsetAsync(dn, 0, 16, s));
  uint32_t *ex0 = T.get<uint32_t>(A.n), *sh = T.get<uint32_t>(B.n), *exb = T.get<uint32_t>(B.n);
  const DView va = view(A), vb = view(B);
  k_mu_keep_a<<<blocks(A.n), 256, 0, s>>>(va, ex0);
  dscan(ex0, ex0, A.n, dn + 1, s, T);
  k_mu_keep_b<<<blocks(B.n), 256, 0, s>>>(va, vb, ex0, sh, exb);
  dscan(exb, exb, B.n, dn + 2, s, T);
  uint32_t hn[4];
  MSCX_HIP(hipMemcpyAsync(hn, dn, 16, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipStreamSynchronize(s));
  const uint32_t n1 = hn[1], n = n1 + hn[2];
  dc.mark("index");
  alloc_nodes(O, (uint32_t)Complex::with_headroom(n), s);
  O.n = n;
  uint32_t *src0 = T.get<uint32_t>(n), *src1 = T.get<uint32_t>(n);
  MSCX_HIP(hipMemsetAsync(src0, 0xFF, (size_t)n * 4, s));
  MSCX_HIP(hipMemsetAsync(src1, 0xFF, (size_t)n * 4, s));
  DView vo = view(O);
  k_mu_nodes<<<blocks((uint64_t)A.n + B.n), 256, 0, s>>>(va, vb, ex0, sh, exb, n1, vo, src0, src1);
  k_map_build<<<blocks(n), 256, 0, s>>>(vo, 0, n, dn);
  k_mu_arcs<false><<<blocks(n), 256, 0, s>>>(va, vb, ex0, sh, src0, src1, vo, rd(ixn));
  dscan(O.cnt, O.off, 2 * (uint64_t)n, dn + 3, s, T);
  const uint32_t words = read_u32(dn + 3, s);
  alloc_arcs(O, words + words / 8 + 1024);
  O.aused = words;
  vo = view(O);
  k_mu_arcs<true><<<blocks(n), 256, 0, s>>>(va, vb, ex0, sh, src0, src1, vo, rd(ixn));
  sort_lists(O, s, T, dn);
  k_mu_pairs<<<blocks(n), 256, 0, s>>>(va, vb, src0, src1, vo, dn);
  check_err(dn, s);
  dc.mark("arcs");
  line_pass(O, ixn, true, s);
  g_dmerges++;
}

void dmerge_down(DComplex &P, DComplex &A, DComplex &B, hipStream_t s) {
  DbgClock dc("merge_down", P.n);
  Rect ixn;
  merge_ixn(A.rect, B.rect, ixn);
  line_pass(P, ixn, false, s);
  dc.mark("line");
  DComplex *kids[2] = {&A, &B};
  for (DComplex *mp : kids) {
    DComplex &m = *mp;
    Temps T(s);
    uint32_t *dn = T.get<uint32_t>(4);  // err, R, arcs
    MSCX_HIP(hipMemsetAsync(dn, 0, 16, s));
    uint32_t *at = T.get<uint32_t>(P.n), *cntj = T.get<uint32_t>(P.n);
    uint8_t *in0 = T.get<uint8_t>(P.n), *proc = T.get<uint8_t>(P.n), *cplx = T.get<uint8_t>(P.n);
    unsigned long long *key = T.get<unsigned long long>(P.n);
    const uint32_t nw = (P.n + 31) / 32;
    uint32_t *bm = T.get<uint32_t>(nw), *woff = T.get<uint32_t>(nw), *ids = T.get<uint32_t>(P.n);
    MSCX_HIP(hipMemsetAsync(bm, 0, (size_t)nw * 4, s));
    const DView vp = view(P);
    k_md_lookup<<<blocks(P.n), 256, 0, s>>>(vp, view(m), at, in0, key, dn);
    k_md_proc<<<blocks(P.n), 256, 0, s>>>(vp, in0, proc, cplx);
    k_md_cmark<<<blocks(P.n), 256, 0, s>>>(vp, bm, dn);
    k_bm_count<<<blocks(nw), 256, 0, s>>>(bm, nw, woff);
    dscan(woff, woff, nw, dn + 3, s, T);
    k_bm_write<<<blocks(nw), 256, 0, s>>>(bm, nw, woff, ids);
    uint32_t hc[4];
    MSCX_HIP(hipMemcpyAsync(hc, dn, 16, hipMemcpyDeviceToHost, s));
    MSCX_HIP(hipStreamSynchronize(s));
    if (hc[0]) throw DFallback("merge_down pair check bits " + std::to_string(hc[0]));
    const uint32_t nc = hc[3];
    uint32_t *cres = T.get<uint32_t>(2 * (uint64_t)nc);
    if (nc) {
      // the reference's walk over the complex components, in index order
      uint32_t *crec = T.get<uint32_t>(4 * (uint64_t)nc);
      k_md_cpack<<<blocks(nc), 256, 0, s>>>(vp, view(m), ids, nc, at, crec);
      std::vector<uint32_t> hid(nc), rec(4 * (size_t)nc), res(2 * (size_t)nc, 0);
      MSCX_HIP(hipMemcpyAsync(hid.data(), ids, (size_t)nc * 4, hipMemcpyDeviceToHost, s));
      MSCX_HIP(hipMemcpyAsync(rec.data(), crec, rec.size() * 4, hipMemcpyDeviceToHost, s));
      MSCX_HIP(hipStreamSynchronize(s));
      std::vector<uint8_t> present(nc), bc(nc);
      for (uint32_t t = 0; t < nc; ++t) {
        present[t] = (uint8_t)rec[4 * (size_t)t + 2];
        bc[t] = present[t] && rec[4 * (size_t)t + 3];
        res[2 * (size_t)t + 1] = NONE;
      }
      for (uint32_t t = 0; t < nc; ++t) {
        if (!(rec[4 * (size_t)t] & F_BC)) continue;
        const uint32_t u = local_of(hid, rec[4 * (size_t)t + 1]);
        if (u == NONE) throw DFallback("merge_down component not closed");
        const bool src_in = present[t], sp_in = present[u];
        if (!src_in && !sp_in) continue;
        res[2 * (size_t)t] |= 1u;
        if (!src_in) {
          present[t] = 1;
          bc[t] = (rec[4 * (size_t)t] & F_BC) != 0;
        }
        if (!sp_in) {
          present[u] = 1;
          bc[u] = (rec[4 * (size_t)u] & F_BC) != 0;
        }
        if (!src_in || !sp_in || !bc[t]) {
          bc[t] = bc[u] = 1;
          res[2 * (size_t)t] |= 2u;
          res[2 * (size_t)u] |= 2u;
          res[2 * (size_t)t + 1] = hid[u];
          res[2 * (size_t)u + 1] = hid[t];
        }
      }
      MSCX_HIP(hipMemcpyAsync(cres, res.data(), res.size() * 4, hipMemcpyHostToDevice, s));
      k_md_cset<<<blocks(nc), 256, 0, s>>>(ids, nc, cres, proc, cplx);
      MSCX_HIP(hipStreamSynchronize(s));  // res is a host vector
    }
    k_md_keys<<<blocks(P.n), 256, 0, s>>>(vp, in0, proc, key, dn);
    k_md_reps<false><<<blocks(P.n), 256, 0, s>>>(vp, in0, proc, key, cntj, nullptr, 0, nullptr, nullptr);
    dscan(cntj, cntj, P.n, dn + 1, s, T);
    uint32_t hn[4];
    MSCX_HIP(hipMemcpyAsync(hn, dn, 16, hipMemcpyDeviceToHost, s));
    MSCX_HIP(hipStreamSynchronize(s));
    if (hn[0]) throw DFallback("merge_down check bits " + std::to_string(hn[0]));
    const uint32_t R = hn[1], n0 = m.n, n = n0 + R;
    if (n > m.ncap) {
      // grow the node arrays (and rebuild the map) before replicas go in
      DComplex g;
      alloc_nodes(g, (uint32_t)Complex::with_headroom(n), s);
      MSCX_HIP(hipMemcpyAsync(g.cell, m.cell, (size_t)n0 * 4, hipMemcpyDeviceToDevice, s));
      MSCX_HIP(hipMemcpyAsync(g.pair, m.pair, (size_t)n0 * 4, hipMemcpyDeviceToDevice, s));
      MSCX_HIP(hipMemcpyAsync(g.fn, m.fn, (size_t)n0 * 4, hipMemcpyDeviceToDevice, s));
      MSCX_HIP(hipMemcpyAsync(g.flags, m.flags, (size_t)n0, hipMemcpyDeviceToDevice, s));
      MSCX_HIP(hipMemcpyAsync(g.off, m.off, (size_t)n0 * 8, hipMemcpyDeviceToDevice, s));
      MSCX_HIP(hipMemcpyAsync(g.cnt, m.cnt, (size_t)n0 * 8, hipMemcpyDeviceToDevice, s));
      g.n = n0;
      k_map_build<<<blocks(n0), 256, 0, s>>>(view(g), 0, n0, dn);
      MSCX_HIP(hipStreamSynchronize(s));
      std::swap(m.node_mem, g.node_mem);
      std::swap(m.map_mem, g.map_mem);
      m.cell = g.cell;
      m.pair = g.pair;
      m.fn = g.fn;
      m.flags = g.flags;
      m.off = g.off;
      m.cnt = g.cnt;
      m.map = g.map;
      m.mmask = g.mmask;
      m.ncap = g.ncap;
    }
    uint32_t *rep_src = T.get<uint32_t>(R);
    k_md_reps<true><<<blocks(P.n), 256, 0, s>>>(vp, in0, proc, key, nullptr, cntj, n0, at, rep_src);
    m.n = n;
    DView vm = view(m);
    k_md_rep_fill<<<blocks(R), 256, 0, s>>>(vp, vm, rep_src, R, n0, dn);
    k_md_pairs<<<blocks(P.n), 256, 0, s>>>(vp, vm, in0, proc, cplx, at);
    if (nc) k_md_pairs_cplx<<<blocks(nc), 256, 0, s>>>(vm, ids, nc, cres, at);
    uint32_t *inv = T.get<uint32_t>(n);
    MSCX_HIP(hipMemsetAsync(inv, 0xFF, (size_t)n * 4, s));
    k_md_inv<<<blocks(P.n), 256, 0, s>>>(P.n, at, inv);
    uint32_t *ncnt = T.get<uint32_t>(2 * (uint64_t)n), *noff = T.get<uint32_t>(2 * (uint64_t)n);
    k_md_arcs<false><<<blocks(n), 256, 0, s>>>(vp, vm, at, inv, ncnt, nullptr, nullptr);
    dscan(ncnt, noff, 2 * (uint64_t)n, dn + 2, s, T);
    const uint32_t words = read_u32(dn + 2, s);
    DComplex na;  // the new pool (only its arc block is used)
    alloc_arcs(na, words + words / 8 + 1024);
    k_md_arcs<true><<<blocks(n), 256, 0, s>>>(vp, vm, at, inv, nullptr, noff, na.arcs);
    MSCX_HIP(hipMemcpyAsync(m.off, noff, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    MSCX_HIP(hipMemcpyAsync(m.cnt, ncnt, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    MSCX_HIP(hipStreamSynchronize(s));
    std::swap(m.arc_mem, na.arc_mem);
    m.arcs = (uint32_t *)m.arc_mem;
    m.acap = na.acap;
    m.aused = words;
    sort_lists(m, s, T, dn);
    check_err(dn, s);
    dc.mark("child");
  }
  g_dmerges++;
}

void dcomplex_copy(const DComplex &a, DComplex &o, hipStream_t s) {
  o.rect = a.rect;
  o.ext = a.ext;
  alloc_nodes(o, a.ncap, s);
  alloc_arcs(o, a.acap);
  o.n = a.n;
  o.aused = a.aused;
  const size_t n = a.n;
  MSCX_HIP(hipMemcpyAsync(o.cell, a.cell, n * 4, hipMemcpyDeviceToDevice, s));
  MSCX_HIP(hipMemcpyAsync(o.pair, a.pair, n * 4, hipMemcpyDeviceToDevice, s));
  MSCX_HIP(hipMemcpyAsync(o.fn, a.fn, n * 4, hipMemcpyDeviceToDevice, s));
  MSCX_HIP(hipMemcpyAsync(o.flags, a.flags, n, hipMemcpyDeviceToDevice, s));
  MSCX_HIP(hipMemcpyAsync(o.off, a.off, n * 8, hipMemcpyDeviceToDevice, s));
  MSCX_HIP(hipMemcpyAsync(o.cnt, a.cnt, n * 8, hipMemcpyDeviceToDevice, s));
  if (a.aused) MSCX_HIP(hipMemcpyAsync(o.arcs, a.arcs, a.aused * 4, hipMemcpyDeviceToDevice, s));
  MSCX_HIP(hipMemcpyAsync(o.map, a.map, ((size_t)a.mmask + 1) * 8, hipMemcpyDeviceToDevice, s));
  MSCX_HIP(hipStreamSynchronize(s));
}

void dcomplex_download(const DComplex &d, Complex &out, hipStream_t s) {
  DbgClock dc("download", d.n);
  const uint32_t n = d.n;
  Temps T(s);
  uint32_t *sz = T.get<uint32_t>(n + 1), *tot = T.get<uint32_t>(1);
  const DView v = view(d);
  k_pack_all<false><<<blocks(n), 256, 0, s>>>(v, sz, nullptr, nullptr);
  dscan(sz, sz, n, tot, s, T);
  const uint32_t words = read_u32(tot, s);
  uint32_t *w = T.get<uint32_t>(words);
  k_pack_all<true><<<blocks(n), 256, 0, s>>>(v, nullptr, sz, w);
  Complex c;
  c.rect = d.rect;
  c.ext = d.ext;
  // the cell map: the device table as it is
  const size_t slots = (size_t)d.mmask + 1;
  MSCX_HIP(hipMemcpyAsync(c.idx.raw_table(slots, n), d.map, slots * 8, hipMemcpyDeviceToHost, s));
  const size_t bytes = ((size_t)words + n) * 4;
  PinLease pin;
  std::unique_ptr<uint32_t[]> big;
  uint32_t *hw;
  if (bytes <= PIN_MAX) {
    hw = (uint32_t *)pin.sc->reserve_host(bytes);
  } else {
    big.reset(new uint32_t[(size_t)words + n]);
    hw = big.get();
  }
  uint32_t *ho = hw + words;
  MSCX_HIP(hipMemcpyAsync(hw, w, (size_t)words * 4, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipMemcpyAsync(ho, sz, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  c.reserve_nodes(Complex::with_headroom(n));
  // on one thread the nodes are appended as they are parsed (every array
  // written once); on several, the arrays are sized first and parsed by ranges
  const int nt = n >= (1u << 16) ? std::min(host_threads(), 16) : 1;
  if (nt == 1) {
    dc.mark("alloc");
    MSCX_HIP(hipStreamSynchronize(s));
    dc.mark("copies");
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t *r = hw + ho[i];
      c.cell.push_back(r[0]);
      float f;
      memcpy(&f, &r[1], 4);
      c.fn.push_back(f);
      c.pair.push_back(r[2]);
      c.flags.push_back((uint8_t)r[3]);
      const uint32_t na = r[4], nd = r[5];
      c.asc.emplace_back();
      c.des.emplace_back();
      ConnSet &a = c.asc.back(), &d = c.des.back();
      for (uint32_t q = 0; q < na; ++q) a.push_sorted(r[6 + q]);
      for (uint32_t q = 0; q < nd; ++q) d.push_sorted(r[6 + na + q]);
    }
    dc.mark("parse");
    out = std::move(c);
    return;
  }
  c.cell.resize(n);
  c.fn.resize(n);
  c.pair.resize(n);
  c.flags.resize(n);
  c.asc.resize(n);
  c.des.resize(n);
  dc.mark("alloc");
  MSCX_HIP(hipStreamSynchronize(s));
  dc.mark("copies");
  auto parse = [&](uint32_t lo, uint32_t hi) {
    for (uint32_t i = lo; i < hi; ++i) {
      const uint32_t *r = hw + ho[i];
      c.cell[i] = r[0];
      memcpy(&c.fn[i], &r[1], 4);
      c.pair[i] = r[2];
      c.flags[i] = (uint8_t)r[3];
      const uint32_t na = r[4], nd = r[5];
      for (uint32_t q = 0; q < na; ++q) c.asc[i].push_sorted(r[6 + q]);
      for (uint32_t q = 0; q < nd; ++q) c.des[i].push_sorted(r[6 + na + q]);
    }
  };
  {
    std::vector<std::thread> th;
    const uint32_t per = (n + nt - 1) / nt;
    for (int t = 1; t < nt; ++t) th.emplace_back(parse, std::min(n, t * per), std::min(n, (t + 1) * per));
    parse(0, std::min(n, per));
    for (auto &x : th) x.join();
  }
  dc.mark("parse");
  out = std::move(c);
}

void dcomplex_upload(const Complex &c, DComplex &o, hipStream_t s) {
  DbgClock dc("upload", c.size());
  const uint32_t n = c.size();
  o.rect = c.rect;
  o.ext = c.ext;
  alloc_nodes(o, (uint32_t)Complex::with_headroom(n), s);
  o.n = n;
  std::vector<uint32_t> cnt(2 * (size_t)n), off(2 * (size_t)n + 1);
  for (uint32_t i = 0; i < n; ++i) {
    cnt[2 * (size_t)i] = c.asc[i].size();
    cnt[2 * (size_t)i + 1] = c.des[i].size();
  }
  off[0] = 0;
  for (size_t k = 0; k < 2 * (size_t)n; ++k) off[k + 1] = off[k] + cnt[k];
  const uint64_t words = off[2 * (size_t)n];
  alloc_arcs(o, words + words / 8 + 1024);
  o.aused = words;
  Temps T(s);
  uint32_t *dfl = T.get<uint32_t>(n), *err = T.get<uint32_t>(1);
  MSCX_HIP(hipMemsetAsync(err, 0, 4, s));
  {
    PinLease pin;
    const size_t bytes = (6 * (size_t)n + words) * 4;
    std::unique_ptr<uint32_t[]> big;
    uint32_t *h;
    if (bytes <= PIN_MAX) {
      h = (uint32_t *)pin.sc->reserve_host(bytes);
    } else {
      big.reset(new uint32_t[6 * (size_t)n + words]);
      h = big.get();
    }
    uint32_t *hc = h, *hp = h + n, *hf = h + 2 * (size_t)n, *hfl = h + 3 * (size_t)n, *hcnt = h + 4 * (size_t)n,
             *ha = h + 6 * (size_t)n;
    auto fill = [&](uint32_t lo, uint32_t hi) {
      for (uint32_t i = lo; i < hi; ++i) {
        hc[i] = c.cell[i];
        hp[i] = c.pair[i];
        memcpy(&hf[i], &c.fn[i], 4);
        hfl[i] = c.flags[i];
        hcnt[2 * (size_t)i] = cnt[2 * (size_t)i];
        hcnt[2 * (size_t)i + 1] = cnt[2 * (size_t)i + 1];
        std::copy(c.asc[i].begin(), c.asc[i].end(), ha + off[2 * (size_t)i]);
        std::copy(c.des[i].begin(), c.des[i].end(), ha + off[2 * (size_t)i + 1]);
      }
    };
    const int nt = n >= (1u << 16) ? std::min(host_threads(), 16) : 1;
    const uint32_t per = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(fill, std::min(n, t * per), std::min(n, (t + 1) * per));
    fill(0, std::min(n, per));
    for (auto &x : th) x.join();
    MSCX_HIP(hipMemcpyAsync(o.cell, hc, (size_t)n * 4, hipMemcpyHostToDevice, s));
    MSCX_HIP(hipMemcpyAsync(o.pair, hp, (size_t)n * 4, hipMemcpyHostToDevice, s));
    MSCX_HIP(hipMemcpyAsync(o.fn, hf, (size_t)n * 4, hipMemcpyHostToDevice, s));
    MSCX_HIP(hipMemcpyAsync(dfl, hfl, (size_t)n * 4, hipMemcpyHostToDevice, s));
    MSCX_HIP(hipMemcpyAsync(o.cnt, hcnt, (size_t)n * 8, hipMemcpyHostToDevice, s));
    MSCX_HIP(hipMemcpyAsync(o.arcs, ha, words * 4, hipMemcpyHostToDevice, s));
    MSCX_HIP(hipStreamSynchronize(s));
  }
  k_flags_from_u32<<<blocks(n), 256, 0, s>>>(dfl, o.flags, n);
  dscan(o.cnt, o.off, 2 * (uint64_t)n, nullptr, s, T);
  k_map_build<<<blocks(n), 256, 0, s>>>(view(o), 0, n, err);
  check_err(err, s);
}

}  // namespace mscx

extern "C" int mscx_merge_stats(uint64_t out[3]) {
  return mscx::guarded([&] {
    MSCX_REQUIRE(out, MSCX_EINVAL, "null argument");
    out[0] = mscx::g_dmerges.load();
    out[1] = mscx::g_dfallbacks.load();
    out[2] = mscx::g_line_passes.load();
  });
}
