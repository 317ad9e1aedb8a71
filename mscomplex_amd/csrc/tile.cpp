// mscx_tile implementation and the batched per-stage launcher.
#include "tile.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

namespace mscx {

static Scratch g_batch;  // descriptors, block prefixes, K array, flags

// Block tile height by tile height (mscx_set_block_rows): the first entry
// whose bound is at least the tile's vertex rows; GRAD_TY otherwise.
static std::mutex g_rows_mu;
static std::vector<std::pair<int32_t, int32_t>> g_rows;

int block_rows_for(int vy) {
  std::lock_guard<std::mutex> lk(g_rows_mu);
  for (const auto &e : g_rows)
    if (vy <= e.first) return e.second;
  return GRAD_TY;
}

void check_rects(const int32_t *rect, const int32_t *ext) {
  for (int k = 0; k < 4; ++k) {
    MSCX_REQUIRE(rect[k] >= 0 && ext[k] >= 0 && rect[k] <= 32766 && ext[k] <= 32766, MSCX_EINVAL,
                 "rect/ext coordinates must lie in [0, 32766]");
    MSCX_REQUIRE((rect[k] & 1) == 0 && (ext[k] & 1) == 0, MSCX_EINVAL,
                 "rect/ext corners must be vertices (even cell coordinates)");
  }
  MSCX_REQUIRE(ext[2] - ext[0] >= 2 && ext[3] - ext[1] >= 2, MSCX_EINVAL, "ext must span at least 2x2 vertices");
  MSCX_REQUIRE(rect[2] - rect[0] >= 2 && rect[3] - rect[1] >= 2, MSCX_EINVAL,
               "rect must span at least 2x2 vertices");
  MSCX_REQUIRE(rect[0] >= ext[0] && rect[1] >= ext[1] && rect[2] <= ext[2] && rect[3] <= ext[3], MSCX_EINVAL,
               "ext must contain rect");
}

// Geometry of a tile over ext and the layout of its one device allocation:
// codes, counts, offsets, scan partials, critical flags, owners, local
// terminals, ring words.  Returns the allocation size; offsets in o[7].
size_t fixed_layout(const int32_t rect[4], const int32_t ext[4], TileDesc &d, size_t o[7]) {
  TileGeom &g = d.g;
  g.vx = (ext[2] - ext[0]) / 2 + 1;
  g.vy = (ext[3] - ext[1]) / 2 + 1;
  g.cw = 2 * g.vx - 1;
  g.ch = 2 * g.vy - 1;
  g.rx0 = rect[0] - ext[0];
  g.ry0 = rect[1] - ext[1];
  g.rx1 = rect[2] - ext[0];
  g.ry1 = rect[3] - ext[1];
  d.ox = ext[0];
  d.oy = ext[1];
  // one collation segment per block tile column: cw = 2 vx - 1 is odd, so
  // ceil(cw / SEG) == ceil(vx / GRAD_TX)
  d.nseg = (g.cw + SEG - 1) / SEG;
  d.pitch = d.nseg * SEG;
  d.own_pitch = d.pitch;
  d.loc_pitch = d.nseg * GRAD_TX;
  const int64_t nrs = (int64_t)g.ch * d.nseg;
  const int64_t nparts = (nrs + SCAN_CHUNK - 1) / SCAN_CHUNK;
  d.ty = block_rows_for(g.vy);
  const int64_t nring = ring_words_max(g, d.nseg);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t at = off;
    off = align_up(off + bytes, 256);
    return at;
  };
  o[0] = take((size_t)g.ch * d.pitch);
  o[1] = take(nrs * 4);
  o[2] = take(nparts * 4 + 4);
  o[3] = take(nrs * 16);
  o[4] = take((size_t)g.ch * d.own_pitch * 4);
  o[5] = take((size_t)g.vy * d.loc_pitch * 4);
  o[6] = take((size_t)nring * 4);
  size_t o_ringidx = take((size_t)nring * 4);
  (void)o_ringidx;
  return off;
}

size_t tile_fixed_bytes(const int32_t rect[4], const int32_t ext[4]) {
  TileDesc d{};
  size_t o[7];
  return fixed_layout(rect, ext, d, o);
}

static void alloc_fixed(mscx_tile *t) {
  TileDesc &d = t->d;
  size_t o[7];
  const size_t bytes = fixed_layout(t->rect, t->ext, d, o);
  t->fixed_bytes = bytes;
  t->fixed = device_pool().alloc(bytes);
  char *base = (char *)t->fixed;
  d.codes = (uint8_t *)(base + o[0]);
  d.seg_offs = (uint32_t *)(base + o[1]);
  d.scan_part = (uint32_t *)(base + o[2]);
  d.critbits = (uint64_t *)(base + o[3]);
  d.own = (uint32_t *)(base + o[4]);
  d.loc = (uint32_t *)(base + o[5]);
  d.ring = (uint32_t *)(base + o[6]);
  d.ringidx = d.ring + ring_words_max(d.g, d.nseg);
}

CritLayout::CritLayout(size_t K) {
  size_t cap = K ? K : 1;
  fn = align_up(cap * 4, 256);
  conn = fn + align_up(cap * 4, 256);
  pair = conn + align_up(cap * 16, 256);
  total = pair + align_up(cap * 4, 256);
}

void run_tiles(mscx_tile *const *tiles, int n) {
  if (n <= 0) return;
  ensure_init();
  Runtime &rt = runtime();
  std::lock_guard<std::mutex> lk(rt.mu);
  hipStream_t s = rt.stream;
  for (int i = 0; i < n; ++i) MSCX_REQUIRE(tiles[i] && tiles[i]->has_input, MSCX_ESTATE, "tile has no input bound");

  // Layout of the batch scratch: descriptors | prefixes (KB_COUNT x (n+1)) |
  // K[n] | error flag (zeroed by the descriptor upload itself).
  const size_t o_desc = 0;
  const size_t o_pref = align_up(sizeof(TileDesc) * n, 256);
  const size_t o_k = align_up(o_pref + sizeof(int) * KB_COUNT * (n + 1), 256);
  const size_t o_err = align_up(o_k + 4 * n, 256);
  const size_t total = o_err + 256;
  char *dev = (char *)g_batch.reserve_dev(total);
  char *host = (char *)g_batch.reserve_host(total);
  TileDesc *hd = (TileDesc *)(host + o_desc);
  int *hp = (int *)(host + o_pref);
  uint32_t *hk = (uint32_t *)(host + o_k);
  uint32_t *herr = (uint32_t *)(host + o_err);
  TileDesc *dd = (TileDesc *)(dev + o_desc);
  int *dp = (int *)(dev + o_pref);
  uint32_t *dk = (uint32_t *)(dev + o_k);
  uint32_t *derr = (uint32_t *)(dev + o_err);

  for (int i = 0; i < n; ++i) {
    hd[i] = tiles[i]->d;
    hd[i].nkrit = dk + i;
    hd[i].kcap = 0;
    hk[i] = 0;
  }
  *herr = 0;
  int nb[KB_COUNT];
  for (int k = 0; k < KB_CRIT; ++k) nb[k] = plan_blocks(k, hd, n, hp + k * (n + 1));

  // descriptor upload, gradient + block-local trace, scan of the critical
  // counts, ring resolution, owner images and the count download need no
  // host decision in between: one HIP graph, built once per launch
  // configuration (staging addresses, sizes, grid sizes) and replayed; the
  // tile descriptors themselves are data that the upload node reads from the
  // pinned staging buffer at every replay.  The graph is built node by node
  // rather than captured from the stream, so other threads' use of the
  // device (including the legacy stream) is never affected.
  StageTimer &tm = stage_timer();
  auto chain = [&] {
    MSCX_HIP(hipMemcpyAsync(dev, host, total, hipMemcpyHostToDevice, s));
    tm.begin(ST_GRADIENT, s);
    launch_grad_local(dd, dp + KB_GRADIENT * (n + 1), n, nb[KB_GRADIENT], derr, s);
    tm.end(ST_GRADIENT, s);
    tm.begin(ST_SCAN, s);
    launch_scan(dd, dp + KB_SCAN * (n + 1), n, nb[KB_SCAN], s);
    tm.end(ST_SCAN, s);
    tm.begin(ST_TRACE, s);
    launch_ring_resolve(dd, dp + KB_RING * (n + 1), n, nb[KB_RING], derr, s);
    tm.end(ST_TRACE, s);
    tm.begin(ST_OWNER, s);
    launch_owner(dd, dp + KB_GRADIENT * (n + 1), n, nb[KB_GRADIENT], s);
    tm.end(ST_OWNER, s);
    MSCX_HIP(hipGetLastError());
    MSCX_HIP(hipMemcpyAsync(hk, dk, o_err + 4 - o_k, hipMemcpyDeviceToHost, s));
  };
  if (rt.graphs) {
    // with stage timing on, the graph records the timer's events around its
    // kernels (a graph of its own: the timing flag is part of the key)
    hipEvent_t ev[8] = {};
    const int evst[4] = {ST_GRADIENT, ST_SCAN, ST_TRACE, ST_OWNER};
    if (tm.on)
      for (int k = 0; k < 4; ++k) {
        tm.ensure(evst[k]);
        ev[2 * k] = tm.a[evst[k]];
        ev[2 * k + 1] = tm.b[evst[k]];
      }
    std::vector<uintptr_t> key{(uintptr_t)n, (uintptr_t)dev, (uintptr_t)host, total, (uintptr_t)tm.on};
    for (int k = 0; k < KB_CRIT; ++k) key.push_back((uintptr_t)nb[k]);
    if (!rt.graph_exec || rt.graph_key != key) {
      if (rt.graph_exec) {
        MSCX_HIP(hipGraphExecDestroy(rt.graph_exec));
        rt.graph_exec = nullptr;
      }
      hipGraph_t g = nullptr;
      MSCX_HIP(build_front_graph(&g, dev, host, total, hk, dk, o_err + 4 - o_k, dd, dp, n, nb, derr,
                                 tm.on ? ev : nullptr));
      const hipError_t e = hipGraphInstantiate(&rt.graph_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      MSCX_HIP(e);
      rt.graph_key = key;
      ++rt.graph_builds;
    } else {
      ++rt.graph_replays;
    }
    MSCX_HIP(hipGraphLaunch(rt.graph_exec, s));
    if (tm.on)
      for (int k = 0; k < 4; ++k) tm.pending[evst[k]] = nb[k == 0 ? KB_GRADIENT : k == 1 ? KB_SCAN
                                                              : k == 2 ? KB_RING : KB_GRADIENT] > 0;
  } else {
    chain();
  }
  MSCX_HIP(hipStreamSynchronize(s));
  tm.collect();
  MSCX_REQUIRE(*herr == 0, MSCX_ELOGIC, "manifold trace did not resolve (invalid gradient)");

  // one block for the critical data of every tile of this launch
  size_t block_bytes = 0;
  std::vector<size_t> coff(n);
  for (int i = 0; i < n; ++i) {
    coff[i] = block_bytes;
    block_bytes += CritLayout(hk[i]).total;
  }
  void *blk = device_pool().alloc(block_bytes);
  std::shared_ptr<void> block(blk, [](void *p) { device_pool().free(p); });
  for (int i = 0; i < n; ++i) {
    CritLayout L(hk[i]);
    char *b = (char *)blk + coff[i];
    tiles[i]->crit_block = block;
    tiles[i]->crit_off = coff[i];
    tiles[i]->crit_block_bytes = block_bytes;
    tiles[i]->d.crit = (uint32_t *)b;
    tiles[i]->d.crit_fn = (float *)(b + L.fn);
    tiles[i]->d.crit_conn = (uint32_t *)(b + L.conn);
    tiles[i]->d.crit_pair = (uint32_t *)(b + L.pair);
    tiles[i]->K = hk[i];
    hd[i].crit = tiles[i]->d.crit;
    hd[i].crit_fn = tiles[i]->d.crit_fn;
    hd[i].crit_conn = tiles[i]->d.crit_conn;
    hd[i].crit_pair = tiles[i]->d.crit_pair;
    hd[i].kcap = hk[i];
  }
  nb[KB_CRIT] = plan_blocks(KB_CRIT, hd, n, hp + KB_CRIT * (n + 1));
  MSCX_HIP(hipMemcpyAsync(dev, host, o_k, hipMemcpyHostToDevice, s));

  tm.begin(ST_COMPACT, s);
  launch_compact(dd, dp + KB_COMPACT * (n + 1), n, nb[KB_COMPACT], s);
  tm.end(ST_COMPACT, s);
  tm.begin(ST_INCIDENCE, s);
  launch_incidence(dd, dp + KB_CRIT * (n + 1), n, nb[KB_CRIT], s);
  tm.end(ST_INCIDENCE, s);
  MSCX_HIP(hipGetLastError());
  MSCX_HIP(hipStreamSynchronize(s));
  tm.collect();
  for (int i = 0; i < n; ++i) tiles[i]->worked = true;
}

void tile_critical(const mscx_tile *t, std::vector<uint32_t> &crit, std::vector<float> &fn,
                   std::vector<uint32_t> &conn4, std::vector<uint32_t> &pair) {
  MSCX_REQUIRE(t->worked, MSCX_ESTATE, "tile has not been worked");
  hipStream_t s = runtime().stream;
  size_t K = t->K;
  crit.resize(K);
  fn.resize(K);
  conn4.resize(4 * K);
  pair.resize(K);
  if (K) {
    MSCX_HIP(hipMemcpyAsync(crit.data(), t->d.crit, K * 4, hipMemcpyDeviceToHost, s));
    MSCX_HIP(hipMemcpyAsync(fn.data(), t->d.crit_fn, K * 4, hipMemcpyDeviceToHost, s));
    MSCX_HIP(hipMemcpyAsync(conn4.data(), t->d.crit_conn, K * 16, hipMemcpyDeviceToHost, s));
    MSCX_HIP(hipMemcpyAsync(pair.data(), t->d.crit_pair, K * 4, hipMemcpyDeviceToHost, s));
  }
  MSCX_HIP(hipStreamSynchronize(s));
}

void tile_codes(const mscx_tile *t, std::vector<uint8_t> &codes) {
  MSCX_REQUIRE(t->worked, MSCX_ESTATE, "tile has not been worked");
  const TileGeom &g = t->d.g;
  codes.resize((size_t)g.cw * g.ch);
  MSCX_HIP(hipMemcpy2DAsync(codes.data(), g.cw, t->d.codes, t->d.pitch, g.cw, g.ch, hipMemcpyDeviceToHost,
                            runtime().stream));
  MSCX_HIP(hipStreamSynchronize(runtime().stream));
}

void tile_owner_index(const mscx_tile *t, std::vector<uint32_t> &own) {
  MSCX_REQUIRE(t->worked, MSCX_ESTATE, "tile has not been worked");
  const TileGeom &g = t->d.g;
  own.resize((size_t)g.cw * g.ch);
  MSCX_HIP(hipMemcpy2DAsync(own.data(), (size_t)g.cw * 4, t->d.own, (size_t)t->d.own_pitch * 4, (size_t)g.cw * 4,
                            g.ch, hipMemcpyDeviceToHost, runtime().stream));
  MSCX_HIP(hipStreamSynchronize(runtime().stream));
}

void tile_owners(const mscx_tile *t, std::vector<uint32_t> &own) {
  tile_owner_index(t, own);
  std::vector<uint32_t> crit(t->K);
  if (t->K) MSCX_HIP(hipMemcpy(crit.data(), t->d.crit, (size_t)t->K * 4, hipMemcpyDeviceToHost));
  for (uint32_t &o : own) {
    if (o == OWN_NONE) continue;
    MSCX_REQUIRE(o < t->K, MSCX_ELOGIC, "owner index out of range");
    o = crit[o];
  }
}

}  // namespace mscx

using namespace mscx;

extern "C" {

int mscx_tile_create(const int32_t rect[4], const int32_t ext[4], mscx_tile **out) {
  return guarded([&] {
    MSCX_REQUIRE(rect && ext && out, MSCX_EINVAL, "null argument");
    check_rects(rect, ext);
    ensure_init();
    mscx_tile *t = new mscx_tile;
    memcpy(t->rect, rect, 16);
    memcpy(t->ext, ext, 16);
    try {
      alloc_fixed(t);
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

int mscx_tile_destroy(mscx_tile *t) {
  return guarded([&] {
    if (!t) return;
    MSCX_HIP(hipStreamSynchronize(runtime().stream));
    device_pool().free(t->fixed);
    device_pool().free(t->verts_owned);
    t->crit_block.reset();
    delete t;
  });
}

int mscx_tile_upload(mscx_tile *t, const float *data, size_t n) {
  return guarded([&] {
    MSCX_REQUIRE(t && data, MSCX_EINVAL, "null argument");
    const TileGeom &g = t->d.g;
    size_t need = (size_t)g.vx * g.vy;
    MSCX_REQUIRE(n == need, MSCX_EINVAL, "input size does not match the ext window");
    for (size_t i = 0; i < n; ++i) MSCX_REQUIRE(!std::isnan(data[i]), MSCX_EINVAL, "input contains NaN");
    if (t->verts_cap < need) {
      device_pool().free(t->verts_owned);
      t->verts_owned = nullptr;
      t->verts_owned = (float *)device_pool().alloc(need * 4);
      t->verts_cap = need;
    }
    MSCX_HIP(hipMemcpyAsync(t->verts_owned, data, need * 4, hipMemcpyHostToDevice, runtime().stream));
    MSCX_HIP(hipStreamSynchronize(runtime().stream));
    t->d.verts = t->verts_owned;
    t->d.vpitch = g.vx;
    t->has_input = true;
    t->worked = false;
  });
}

int mscx_tile_bind_device(mscx_tile *t, const float *dev_ptr, int64_t row_pitch) {
  return guarded([&] {
    MSCX_REQUIRE(t && dev_ptr, MSCX_EINVAL, "null argument");
    MSCX_REQUIRE(row_pitch >= t->d.g.vx, MSCX_EINVAL, "row pitch smaller than the window width");
    t->d.verts = dev_ptr;
    t->d.vpitch = row_pitch;
    t->has_input = true;
    t->worked = false;
  });
}

int mscx_tile_work(mscx_tile *t) {
  return guarded([&] {
    MSCX_REQUIRE(t, MSCX_EINVAL, "null tile");
    run_tiles(&t, 1);
  });
}

int mscx_tiles_work(mscx_tile *const *tiles, int n) {
  return guarded([&] {
    MSCX_REQUIRE(tiles && n >= 0, MSCX_EINVAL, "bad tile array");
    run_tiles(tiles, n);
  });
}

int mscx_tile_num_critical(const mscx_tile *t, uint32_t *K) {
  return guarded([&] {
    MSCX_REQUIRE(t && K, MSCX_EINVAL, "null argument");
    MSCX_REQUIRE(t->worked, MSCX_ESTATE, "tile has not been worked");
    *K = t->K;
  });
}

int mscx_tile_download_codes(const mscx_tile *t, uint8_t *codes) {
  return guarded([&] {
    MSCX_REQUIRE(t && codes, MSCX_EINVAL, "null argument");
    std::vector<uint8_t> c;
    tile_codes(t, c);
    memcpy(codes, c.data(), c.size());
  });
}

int mscx_tile_download(const mscx_tile *t, int8_t *flags, uint32_t *pairs, uint32_t *own, uint32_t *crit) {
  return guarded([&] {
    MSCX_REQUIRE(t, MSCX_EINVAL, "null tile");
    const TileGeom &g = t->d.g;
    if (flags || pairs) {
      std::vector<uint8_t> c;
      tile_codes(t, c);
      for (int y = 0; y < g.ch; ++y)
        for (int x = 0; x < g.cw; ++x) {
          size_t i = (size_t)y * g.cw + x;
          uint8_t code = c[i];
          if (flags) flags[i] = (int8_t)(code & 3);
          if (pairs) {
            int d = code_dir(code);
            pairs[i] = d ? ((uint32_t)(uint16_t)(t->ext[0] + x + dir_dx(d)) |
                            ((uint32_t)(uint16_t)(t->ext[1] + y + dir_dy(d)) << 16))
                         : 0xFFFFFFFFu;
          }
        }
    }
    if (own) {
      std::vector<uint32_t> o;
      tile_owners(t, o);
      memcpy(own, o.data(), o.size() * 4);
    }
    if (crit && t->K) {
      MSCX_REQUIRE(t->worked, MSCX_ESTATE, "tile has not been worked");
      MSCX_HIP(hipMemcpy(crit, t->d.crit, (size_t)t->K * 4, hipMemcpyDeviceToHost));
    }
  });
}

int mscx_tile_download_critical(const mscx_tile *t, float *fn, uint32_t *conn4, uint32_t *pair_idx) {
  return guarded([&] {
    MSCX_REQUIRE(t, MSCX_EINVAL, "null tile");
    std::vector<uint32_t> c, cn, p;
    std::vector<float> f;
    tile_critical(t, c, f, cn, p);
    if (fn) memcpy(fn, f.data(), f.size() * 4);
    if (conn4) memcpy(conn4, cn.data(), cn.size() * 4);
    if (pair_idx) memcpy(pair_idx, p.data(), p.size() * 4);
  });
}

int mscx_set_block_rows(int32_t n, const int32_t *max_rows, const int32_t *rows) {
  return mscx::guarded([&] {
    MSCX_REQUIRE(n >= 0 && n <= 16, MSCX_EINVAL, "at most 16 block-height classes");
    MSCX_REQUIRE(n == 0 || (max_rows && rows), MSCX_EINVAL, "null class array");
    std::vector<std::pair<int32_t, int32_t>> t;
    for (int32_t k = 0; k < n; ++k) {
      MSCX_REQUIRE(rows[k] == mscx::GRAD_TY || rows[k] == mscx::GRAD_TY_LOW, MSCX_EINVAL,
                   "block height must be " + std::to_string(mscx::GRAD_TY) + " or " +
                       std::to_string(mscx::GRAD_TY_LOW) + " rows");
      MSCX_REQUIRE(k == 0 || max_rows[k] > max_rows[k - 1], MSCX_EINVAL, "class bounds must increase");
      t.emplace_back(max_rows[k], rows[k]);
    }
    std::lock_guard<std::mutex> lk(mscx::g_rows_mu);
    mscx::g_rows.swap(t);
  });
}

}  // extern "C"
