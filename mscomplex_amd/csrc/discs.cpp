// postMergeFillDiscs for many leaves at once: the critical-point pass and
// the saddle tracing run on host threads; the per-vertex / per-quad label
// pass (the bulk of the work: one owner lookup, redirects and an extremum
// map search per interior cell) runs on the GPU over every leaf in one
// launch, reading the owner images that are already resident in HBM.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "complex.hpp"
#include "devsort.hpp"
#include "dmerge.hpp"
#include <chrono>

#include "runtime.hpp"
#include "tile.hpp"

namespace mscx {

// Staging of one call: host codes + grouped disc cells (disc), host staging of
// descriptors and tables (stage); the device descriptors, tables and labels
// are a block of the device pool for the duration of the call.  Calls
// from different threads (grids finishing together in a batch or on the
// serving loop's workers) each lease their own set, so their host work runs
// concurrently; a set is kept for reuse once returned, so steady-state calls
// allocate nothing.
// disc slots traced on the GPU and on the host, and calls whose GPU lists
// were dropped because the chains were no gradient's V-paths
static std::atomic<uint64_t> g_trace_gpu{0}, g_trace_host{0}, g_trace_dropped{0}, g_trace_wide{0}, g_trace_longest{0};

static double now_wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct DiscScratch {
  Scratch codes, disc, stage, trace, lists;
};

// A block of the device pool that goes back once `s` has finished with it.
struct PoolBlock {
  void *p;
  hipStream_t s;
  ~PoolBlock() {
    (void)hipStreamSynchronize(s);
    device_pool().free(p);
  }
};
static std::mutex g_pool_mu;
static std::vector<std::unique_ptr<DiscScratch>> g_pool;
struct DiscLease {
  std::unique_ptr<DiscScratch> s;
  DiscLease() {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (g_pool.empty()) {
      s.reset(new DiscScratch);
    } else {
      s = std::move(g_pool.back());
      g_pool.pop_back();
    }
  }
  ~DiscLease() {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool.push_back(std::move(s));
  }
  DiscLease(const DiscLease &) = delete;
  DiscLease &operator=(const DiscLease &) = delete;
};

// Byte offsets of the tiles' gradient codes in one host staging buffer.
static std::vector<size_t> codes_offsets(mscx_tile *const *tiles, int n) {
  std::vector<size_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + align_up((size_t)tiles[i]->d.g.ch * tiles[i]->d.pitch, 256);
  return off;
}

// The gradient codes of a set of worked tiles, copied to pinned host memory
// on the auxiliary stream ahead of their manifold pass (the copies overlap
// the host merge tree); the lease that holds the staging goes back to the
// pool once the copies have finished.
struct CodesStage {
  DiscLease lease;
  std::vector<mscx_tile *> tiles;
  std::vector<size_t> off;
  const uint8_t *host = nullptr;
  hipEvent_t ev = nullptr;
  ~CodesStage() {
    if (ev) {
      (void)hipEventSynchronize(ev);
      (void)hipEventDestroy(ev);
    }
  }
};

std::shared_ptr<CodesStage> stage_codes(mscx_tile *const *tiles, int n) {
  if (n <= 0) return nullptr;
  ensure_init();
  for (int i = 0; i < n; ++i) MSCX_REQUIRE(tiles[i]->worked, MSCX_ESTATE, "tile has not been worked");
  std::shared_ptr<CodesStage> st(new CodesStage);
  st->tiles.assign(tiles, tiles + n);
  st->off = codes_offsets(tiles, n);
  char *h = (char *)st->lease.s->codes.reserve_host(st->off[n]);
  hipStream_t s = runtime().aux;
  for (int i = 0; i < n; ++i)
    MSCX_HIP(hipMemcpyAsync(h + st->off[i], tiles[i]->d.codes, (size_t)tiles[i]->d.g.ch * tiles[i]->d.pitch,
                            hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipEventCreateWithFlags(&st->ev, hipEventDisableTiming));
  MSCX_HIP(hipEventRecord(st->ev, s));
  st->host = (const uint8_t *)h;
  return st;
}

// Work items of a per-leaf loop split into chunks of at most `chunk`
// elements: (leaf, begin, end), the largest leaves' chunks first, so that a
// few large leaves spread over the threads and small ones fill the gaps.
struct Chunk {
  int leaf;
  size_t begin, end;
};
static std::vector<Chunk> chunks_of(const std::vector<size_t> &count, size_t chunk) {
  std::vector<int> order(count.size());
  for (size_t i = 0; i < count.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return count[a] > count[b]; });
  std::vector<Chunk> out;
  for (int i : order)
    for (size_t b = 0; b < std::max<size_t>(count[i], 1); b += chunk) out.push_back({i, b, std::min(count[i], b + chunk)});
  return out;
}

// The label of disc_labels_host (grid_dataset.cpp:1142-1187) depends on a
// cell only through its owner o and whether the cell is itself critical:
// label = map[o2] with o2 = own2(o) when o lies in the rect, where own2(c)
// is c's redirect if it has one and its owner otherwise.  A critical cell
// owns itself and is looked up through own2 first.  Both cases are tabled
// per critical index k of the leaf: owned[k] (cell not critical, owner k)
// and self[k] (the critical cell k itself).  A cell owning another cell is
// always critical in the leaf, except possibly a redirect target; for such a
// target the owner is read from the device image.
// Entries [kb, ke) of the tables (the redirect check runs with the chunk
// that starts at 0).
static void disc_label_tables(const Complex &m, uint32_t K, const Rect &rect, const DiscPrep &dp,
                              const mscx_tile *t, uint32_t *owned, uint32_t *self, uint32_t kb, uint32_t ke) {
  // Every cell the reference's label() looks at here is a node's cell (a
  // critical point of the leaf, a redirected saddle or its pair, or the
  // extremum a redirect points to), so its lookups in the cell map and the
  // redirect map are taken by node index.
  auto in_r = [&](uint32_t c) { return rect.contains(cx_of(c), cy_of(c)); };
  // own[] of node j's cell inside the ext, as a node index: a critical point
  // of the leaf owns itself; any other node's cell is read from the owner
  // image (UINT32_MAX: no owner)
  auto owner_node = [&](uint32_t j) -> uint32_t {
    if (j < K) return j;
    const uint32_t c = m.cell[j];
    uint32_t oi = 0;
    const int64_t off = (int64_t)(cy_of(c) - t->ext[1]) * t->d.own_pitch + (cx_of(c) - t->ext[0]);
    MSCX_HIP(hipMemcpy(&oi, t->d.own + off, 4, hipMemcpyDeviceToHost));
    return oi == OWN_NONE ? UINT32_MAX : oi;
  };
  // own2 of node j's cell (one redirect, else the owner)
  auto own2 = [&](uint32_t j) {
    const uint32_t r = dp.redir_node[j];
    return r != UINT32_MAX ? r : owner_node(j);
  };
  // label() of the reference on node j's cell: the surviving extremum (of
  // dimension edim) that owns it after one redirect, LAB_ERR if the owner is
  // no surviving extremum, UINT32_MAX if the cell has no owner
  auto label = [&](uint32_t j, uint32_t edim) -> uint32_t {
    if (in_r(m.cell[j])) j = own2(j);
    if (j == UINT32_MAX) return UINT32_MAX;
    return (cell_dim(m.cell[j]) == edim && !(m.flags[j] & CPF_BCANC)) ? j : LAB_ERR;
  };
  // a redirect at a vertex or quad that the leaf does not hold as critical
  // would change the label of a non-critical cell, which the tables cannot
  // express
  if (kb == 0)
    for (uint32_t j = K; j < m.size(); ++j)
      if (dp.redir_node[j] != UINT32_MAX && cell_dim(m.cell[j]) != 1 && in_r(m.cell[j]))
        throw std::logic_error("owner redirect at a cell that is not critical in the leaf");
  for (uint32_t k = kb; k < ke; ++k) {
    const uint32_t dim = cell_dim(m.cell[k]);
    if (dim == 1) {  // saddles own no vertex or quad
      owned[k] = self[k] = LAB_ERR;
      continue;
    }
    owned[k] = label(k, dim);
    const uint32_t r = dp.redir_node[k];
    self[k] = label(r != UINT32_MAX ? r : k, dim);
  }
}

void fill_discs_device(mscx_tile *const *tiles, Complex *const *cs, int n, CodesStage *pre) {
  if (n <= 0) return;
  ensure_init();
  // The label pass runs on the auxiliary stream, so it overlaps the tile
  // stages of other grids on the work stream; the owner images and codes it
  // reads were completed (and synchronised) by run_tiles.  Concurrent calls
  // share the stream (a synchronisation may also wait for another call's
  // copies) but not their staging.  Codes staged ahead for exactly these
  // tiles are used as they are (with the staging's lease).
  const bool staged = pre && pre->tiles.size() == (size_t)n && std::equal(pre->tiles.begin(), pre->tiles.end(), tiles);
  std::unique_ptr<DiscLease> own_lease;
  if (!staged) own_lease.reset(new DiscLease);
  DiscLease &lease = staged ? pre->lease : *own_lease;
  Scratch &g_disc = lease.s->disc, &g_stage = lease.s->stage;
  hipStream_t s = runtime().aux;
  for (int i = 0; i < n; ++i) MSCX_REQUIRE(tiles[i]->worked, MSCX_ESTATE, "tile has not been worked");

  std::vector<Rect> rect(n), ext(n);
  std::vector<size_t> nv(n), nq(n);
  const std::vector<size_t> code_off = codes_offsets(tiles, n);
  for (int i = 0; i < n; ++i) {
    const mscx_tile *t = tiles[i];
    rect[i] = Rect{t->rect[0], t->rect[1], t->rect[2], t->rect[3]};
    ext[i] = Rect{t->ext[0], t->ext[1], t->ext[2], t->ext[3]};
    size_t nvx = (rect[i].x1 - rect[i].x0) / 2 + 1, nvy = (rect[i].y1 - rect[i].y0) / 2 + 1;
    nv[i] = nvx * nvy;
    nq[i] = (nvx - 1) * (nvy - 1);
  }

  // 1. gradient codes to pinned host memory (needed by the critical-point
  //    pass and the saddle tracing)
  StageTimer &tm = stage_timer();
  double tp = now_wall();
  std::vector<DiscPrep> dp(n);
  size_t lab_total = 0, nelem = 0, nlab = 0;
  std::vector<uint32_t> lab_base(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    lab_total += align_up((nv[i] + nq[i]) * 4, 256);
    nelem += nv[i] + nq[i];
    lab_base[i + 1] = lab_base[i] + cs[i]->size();
  }
  nlab = lab_base[n];
  MSCX_REQUIRE(nlab < 0xFFFFFFFFull && nelem < (1ull << 40), MSCX_EINVAL, "too many cells for one manifold pass");
  const uint8_t *codes_h;
  if (staged) {
    MSCX_HIP(hipEventSynchronize(pre->ev));
    codes_h = pre->host;
  } else {
    char *h = (char *)lease.s->codes.reserve_host(code_off[n]);
    for (int i = 0; i < n; ++i)
      MSCX_HIP(hipMemcpyAsync(h + code_off[i], tiles[i]->d.codes, (size_t)tiles[i]->d.g.ch * tiles[i]->d.pitch,
                              hipMemcpyDeviceToHost, s));
    MSCX_HIP(hipStreamSynchronize(s));
    codes_h = (const uint8_t *)h;
  }
  // host: grouped cells (one word per vertex and quad) | disc offsets
  const size_t h_cells = 0, h_off = align_up(4 * nelem, 256);
  char *host = (char *)g_disc.reserve_host(h_off + 4 * (nlab + 1));
  std::vector<CodesView> cv(n);
  for (int i = 0; i < n; ++i) cv[i] = CodesView{codes_h + code_off[i], tiles[i]->d.pitch, ext[i].x0, ext[i].y0};

  tm.add_host(ST_DISC_CODES, now_wall() - tp);
  tp = now_wall();
  // 2. critical-point pass (seeds, redirects, surviving maps)
  std::vector<size_t> nk(n);
  for (int i = 0; i < n; ++i) nk[i] = tiles[i]->K;
  {
    const std::vector<Chunk> whole = chunks_of(nk, ~(size_t)0);  // one item per leaf, largest first
    parallel_for((int)whole.size(), [&](int q) {
      const int i = whole[q].leaf;
      disc_prepare(*cs[i], rect[i], cv[i], dp[i], false);
    });
  }

  tm.add_host(ST_DISC_PREP, now_wall() - tp);
  // the event after the trace counts
  struct OneEvent {
    hipEvent_t e = nullptr;
    OneEvent() { MSCX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
    ~OneEvent() { (void)hipEventDestroy(e); }
  } ev_cnt;
  // 3. disc tracing on the GPU (the saddles' 1-manifolds), on a stream of
  //    its own so that its latency-bound passes overlap the label pass and
  //    the grouping: every slot's list is counted first, while the host
  //    builds the label tables; the counts are scanned into each complex's
  //    offsets once the labels and their grouping are under way, and the
  //    lists are written after them
  StreamLease trace_lease;
  hipStream_t ts = trace_lease.s;
  tp = now_wall();
  std::vector<size_t> nslot(n), slot_base(n + 1, 0), init_base(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    nslot[i] = 2 * (size_t)cs[i]->size();
    slot_base[i + 1] = slot_base[i] + nslot[i];
    init_base[i + 1] = init_base[i] + dp[i].init_cells.size();
  }
  const size_t S = slot_base[n], NI = init_base[n];
  // device: descriptors | prefix | error | init_off (S + n) | traced (S) |
  // init cells (NI) | counts (S) | offsets (S); the lists follow in their
  // own block once their total is known
  const size_t t_desc = 0, t_pref = align_up(sizeof(DiscTraceDesc) * n, 256);
  const size_t t_err = align_up(t_pref + 4 * (n + 1), 256);
  const size_t t_ioff = t_err + 256;
  const size_t t_trc = align_up(t_ioff + 4 * (S + n), 256);
  const size_t t_init = align_up(t_trc + S, 256);
  const size_t t_cnt = align_up(t_init + 4 * NI, 256);
  const size_t t_off = align_up(t_cnt + 4 * S, 256);
  const size_t t_bytes = align_up(t_off + 4 * S, 256);
  PoolBlock trblk{device_pool().alloc(t_bytes), ts};
  char *tdev = (char *)trblk.p;
  char *th = (char *)lease.s->trace.reserve_host(t_bytes);
  DiscTraceDesc *td = (DiscTraceDesc *)(th + t_desc);
  uint32_t *terr_h = (uint32_t *)(th + t_err);
  parallel_for(n, [&](int i) {
    const DiscPrep &P = dp[i];
    memcpy(th + t_ioff + 4 * (slot_base[i] + i), P.init_off.data(), 4 * (nslot[i] + 1));
    memcpy(th + t_trc + slot_base[i], P.traced.data(), nslot[i]);
    if (!P.init_cells.empty()) memcpy(th + t_init + 4 * init_base[i], P.init_cells.data(), 4 * P.init_cells.size());
  });
  for (int i = 0; i < n; ++i) {
    const mscx_tile *t = tiles[i];
    DiscTraceDesc &D = td[i];
    D.codes = t->d.codes;
    D.pitch = t->d.pitch;
    D.ox = ext[i].x0;
    D.oy = ext[i].y0;
    D.ex1 = ext[i].x1;
    D.ey1 = ext[i].y1;
    D.rx0 = rect[i].x0;
    D.ry0 = rect[i].y0;
    D.rx1 = rect[i].x1;
    D.ry1 = rect[i].y1;
    D.init_off = (const uint32_t *)(tdev + t_ioff) + slot_base[i] + i;
    D.init_cells = (const uint32_t *)(tdev + t_init) + init_base[i];
    D.traced = (const uint8_t *)(tdev + t_trc) + slot_base[i];
    D.nslot = (uint32_t)nslot[i];
    D.count = (uint32_t *)(tdev + t_cnt) + slot_base[i];
    D.out_off = (const uint32_t *)(tdev + t_off) + slot_base[i];
    D.out = nullptr;
    D.err = (uint32_t *)(tdev + t_err);
  }
  const int tblocks = plan_disc_trace_blocks(td, n, (int *)(th + t_pref));
  *terr_h = 0;
  MSCX_HIP(hipMemcpyAsync(tdev, th, t_cnt, hipMemcpyHostToDevice, ts));
  launch_disc_trace((const DiscTraceDesc *)tdev, (const int *)(tdev + t_pref), n, tblocks, false, ts);
  MSCX_HIP(hipGetLastError());
  // counts (and the error word) back
  uint32_t *cnt_h = (uint32_t *)(th + t_cnt);
  MSCX_HIP(hipMemcpyAsync(cnt_h, tdev + t_cnt, 4 * S, hipMemcpyDeviceToHost, ts));
  MSCX_HIP(hipMemcpyAsync(terr_h, tdev + t_err, 4, hipMemcpyDeviceToHost, ts));
  MSCX_HIP(hipEventRecord(ev_cnt.e, ts));
  double paths_s = now_wall() - tp;
  tp = now_wall();
  // 4. label tables per critical index (host) + label pass (device)
  std::vector<size_t> tab_off(n + 1, 0);
  for (int i = 0; i < n; ++i) tab_off[i + 1] = tab_off[i] + align_up(2 * (size_t)tiles[i]->K * 4, 256);
  const size_t o_desc = 0, o_pref = align_up(sizeof(DiscDesc) * n, 256);
  const size_t o_err = align_up(o_pref + 4 * (n + 1), 256);
  const size_t o_tab = o_err + 256;
  const size_t o_lab = o_tab + tab_off[n];
  const size_t dev_bytes = o_lab + lab_total;
  // device tables and labels come from the library's pool (cached between
  // calls, counted against a server's HBM budget while in use); the block
  // goes back once the auxiliary stream has finished with it
  PoolBlock devblk{device_pool().alloc(dev_bytes), s};
  char *dev = (char *)devblk.p;
  char *stage = (char *)g_stage.reserve_host(o_lab);
  DiscDesc *hd = (DiscDesc *)(stage + o_desc);
  {
    // entries of a leaf's tables are independent: large leaves are split
    const std::vector<Chunk> items = chunks_of(nk, 1 << 16);
    parallel_for((int)items.size(), [&](int q) {
      const int i = items[q].leaf;
      const uint32_t K = tiles[i]->K;
      uint32_t *owned = (uint32_t *)(stage + o_tab + tab_off[i]);
      disc_label_tables(*cs[i], K, rect[i], dp[i], tiles[i], owned, owned + K, (uint32_t)items[q].begin,
                        (uint32_t)items[q].end);
    });
  }
  tm.add_host(ST_DISC_TABLES, now_wall() - tp);
  size_t lo = 0;
  for (int i = 0; i < n; ++i) {
    const mscx_tile *t = tiles[i];
    DiscDesc &D = hd[i];
    D.own = t->d.own;
    D.own_pitch = t->d.own_pitch;
    D.codes = t->d.codes;
    D.pitch = t->d.pitch;
    D.ox = t->ext[0];
    D.oy = t->ext[1];
    D.rx0 = rect[i].x0;
    D.ry0 = rect[i].y0;
    D.rx1 = rect[i].x1;
    D.ry1 = rect[i].y1;
    D.lab_owned = (const uint32_t *)(dev + o_tab + tab_off[i]);
    D.lab_self = D.lab_owned + t->K;
    D.nlab = t->K;
    D.lab_v = (uint32_t *)(dev + o_lab + lo);
    D.lab_q = D.lab_v + nv[i];
    lo += align_up((nv[i] + nq[i]) * 4, 256);
    D.err = (uint32_t *)(dev + o_err);
    D.lab_base = lab_base[i];
    D.elem_base = 0;
  }
  {
    size_t eb = 0;
    for (int i = 0; i < n; ++i) {
      hd[i].elem_base = eb;
      eb += nv[i] + nq[i];
    }
  }
  int nblocks = plan_disc_blocks(hd, n, (int *)(stage + o_pref));
  *(uint32_t *)(stage + o_err) = 0;
  MSCX_HIP(hipMemcpyAsync(dev, stage, o_lab, hipMemcpyHostToDevice, s));
  launch_disc_labels((const DiscDesc *)dev, (const int *)(dev + o_pref), n, nblocks, s);
  MSCX_HIP(hipGetLastError());
  // 5. grouping of the extremum discs on the GPU: the rect's vertices and
  //    quads sorted by label (a stable radix sort of (label, cell) pairs
  //    whose input is in raster order), the discs' offsets found in the
  //    sorted labels; both come back in one copy each
  const uint32_t sentinel = (uint32_t)nlab;
  int nbits = 0;
  while (nbits < 32 && ((uint64_t)sentinel >> nbits)) ++nbits;
  const size_t g_tmp = sort_temp_words(nelem);
  const size_t go_keys = 0, go_vals = align_up(4 * nelem, 256), go_keys2 = 2 * go_vals, go_vals2 = 3 * go_vals;
  const size_t go_off = 4 * go_vals;
  const size_t go_tmp = align_up(go_off + 4 * (nlab + 1), 256), go_pref = align_up(go_tmp + 4 * g_tmp, 256);
  const size_t g_bytes = go_pref + 4 * (n + 1);
  PoolBlock grpblk{device_pool().alloc(g_bytes), s};
  char *gdev = (char *)grpblk.p;
  std::vector<int> gpref(n + 1);
  const int gblocks = plan_disc_group_blocks(hd, n, gpref.data());
  MSCX_HIP(hipMemcpyAsync(gdev + go_pref, gpref.data(), 4 * (n + 1), hipMemcpyHostToDevice, s));
  launch_disc_group_keys((const DiscDesc *)dev, (const int *)(gdev + go_pref), n, gblocks, sentinel,
                         (uint32_t *)(gdev + go_keys), (uint32_t *)(gdev + go_vals), s);
  MSCX_HIP(hipGetLastError());
  const bool alt = radix_sort_pairs((uint32_t *)(gdev + go_keys), (uint32_t *)(gdev + go_vals),
                                    (uint32_t *)(gdev + go_keys2), (uint32_t *)(gdev + go_vals2), nelem, nbits,
                                    (uint32_t *)(gdev + go_tmp), s);
  launch_disc_group_offsets((const uint32_t *)(gdev + (alt ? go_keys2 : go_keys)), nelem, sentinel,
                            (uint32_t *)(gdev + go_off), s);
  MSCX_HIP(hipGetLastError());
  uint32_t *hcells = (uint32_t *)(host + h_cells), *hoff = (uint32_t *)(host + h_off);
  MSCX_HIP(hipMemcpyAsync(hoff, gdev + go_off, 4 * (nlab + 1), hipMemcpyDeviceToHost, s));
  // only the labelled cells (the sentinel sorts last); their count is the
  // last offset, so the copy takes every cell and the host uses the front
  MSCX_HIP(hipMemcpyAsync(hcells, gdev + (alt ? go_vals2 : go_vals), 4 * nelem, hipMemcpyDeviceToHost, s));
  OneEvent ev_grp;
  MSCX_HIP(hipEventRecord(ev_grp.e, s));
  uint32_t *herr = (uint32_t *)(stage + o_err);
  MSCX_HIP(hipMemcpyAsync(herr, dev + o_err, 4, hipMemcpyDeviceToHost, s));

  tp = now_wall();
  MSCX_HIP(hipEventSynchronize(ev_cnt.e));
  const bool gpu_trace = *terr_h == 0;
  std::vector<size_t> out_base(n + 1, 0);
  // slots with more chains than a wave follows are traced here: their
  // lists, by leaf, in slot order
  std::vector<std::vector<std::pair<uint32_t, std::vector<uint32_t>>>> host_slots(n);
  if (gpu_trace) {
    // each complex's offsets by slot; the device offsets are the same per leaf
    uint32_t *off_h = (uint32_t *)(th + t_off);
    std::atomic<uint64_t> wide(0), longest(0);
    parallel_for(n, [&](int i) {
      std::vector<uint32_t> &o = cs[i]->sd_off;
      o.assign(nslot[i] + 1, 0);
      const uint32_t *c = cnt_h + slot_base[i];
      const DiscPrep &P = dp[i];
      uint64_t nw = 0, lmax = 0;
      for (size_t k = 0; k < nslot[i]; ++k) {
        uint32_t len = c[k];
        if (len == DISC_TRACE_WIDE) throw std::logic_error("disc slot left unmeasured by the wide tracing");
        if (len != DISC_TRACE_HOST && len >= DISC_TRACE_WIDE_LEN) {
          len -= DISC_TRACE_WIDE_LEN;
          ++nw;
        }
        if (len != DISC_TRACE_HOST) lmax = std::max<uint64_t>(lmax, len);
        if (len == DISC_TRACE_HOST) {
          std::vector<uint32_t> lst;
          disc_trace_one(P.init_cells.data() + P.init_off[k], P.init_off[k + 1] - P.init_off[k], (int)(k & 1),
                         rect[i], ext[i], cv[i], lst);
          len = (uint32_t)lst.size();
          host_slots[i].emplace_back((uint32_t)k, std::move(lst));
        }
        o[k + 1] = o[k] + len;
      }
      memcpy(off_h + slot_base[i], o.data(), 4 * nslot[i]);
      wide += nw;
      for (uint64_t m = longest.load(); lmax > m && !longest.compare_exchange_weak(m, lmax);) {
      }
    });
    g_trace_wide += wide.load();
    for (uint64_t m = g_trace_longest.load(); longest.load() > m && !g_trace_longest.compare_exchange_weak(m, longest.load());) {
    }
    for (int i = 0; i < n; ++i) out_base[i + 1] = out_base[i] + cs[i]->sd_off[nslot[i]];
    size_t nh = 0;
    for (int i = 0; i < n; ++i) nh += host_slots[i].size();
    g_trace_gpu += S - nh;
    g_trace_host += nh;
  } else {
    g_trace_dropped++;
    g_trace_host += S;
  }
  const size_t NO = out_base[n];
  PoolBlock outblk{device_pool().alloc(std::max<size_t>(4 * NO, 256)), ts};
  uint32_t *out_h = nullptr;
  if (gpu_trace) {
    for (int i = 0; i < n; ++i) td[i].out = (uint32_t *)outblk.p + out_base[i];
    MSCX_HIP(hipMemcpyAsync(tdev, th, t_err, hipMemcpyHostToDevice, ts));
    MSCX_HIP(hipMemcpyAsync(tdev + t_off, th + t_off, 4 * S, hipMemcpyHostToDevice, ts));
    launch_disc_trace((const DiscTraceDesc *)tdev, (const int *)(tdev + t_pref), n, tblocks, true, ts);
    MSCX_HIP(hipGetLastError());
    out_h = (uint32_t *)lease.s->lists.reserve_host(std::max<size_t>(4 * NO, 4));
    if (NO) MSCX_HIP(hipMemcpyAsync(out_h, outblk.p, 4 * NO, hipMemcpyDeviceToHost, ts));
  }
  tm.add_host(ST_DISC_PATHS, paths_s + now_wall() - tp);

  // 6. the extremum discs into the complexes as the grouping's copies land;
  //    without the GPU lists (the chains were no gradient's V-paths) every
  //    leaf's saddle discs are traced here instead
  tp = now_wall();
  std::atomic<int64_t> wait_ns(0), trace_ns(0), group_ns(0);
  try {
    parallel_for(2 * n, [&](int q) {
      const double w0 = now_wall();
      if (q >= n) {
        if (gpu_trace) return;
        const int i = q - n;
        disc_trace_host(*cs[i], dp[i], rect[i], ext[i], cv[i]);
        trace_ns += (int64_t)(1e9 * (now_wall() - w0));
        return;
      }
      const int i = q;
      MSCX_HIP(hipEventSynchronize(ev_grp.e));
      const double w1 = now_wall();
      wait_ns += (int64_t)(1e9 * (w1 - w0));
      Complex &m = *cs[i];
      const uint32_t b0 = lab_base[i], nn = m.size();
      m.xd_off.resize((size_t)nn + 1);
      for (uint32_t k = 0; k <= nn; ++k) m.xd_off[k] = hoff[b0 + k] - hoff[b0];
      const size_t nc = hoff[b0 + nn] - hoff[b0];
      if (nc > m.xd_cells.capacity()) {
        std::vector<uint32_t>().swap(m.xd_cells);
        m.xd_cells.reserve(nc);
        advise_huge(m.xd_cells.data(), 4 * nc);
      }
      m.xd_cells.assign(hcells + hoff[b0], hcells + hoff[b0 + nn]);
      group_ns += (int64_t)(1e9 * (now_wall() - w1));
    });
  } catch (...) {
    // no copy may still be landing in the staging buffers
    (void)hipStreamSynchronize(s);
    (void)hipStreamSynchronize(ts);
    throw;
  }
  MSCX_HIP(hipStreamSynchronize(s));
  MSCX_HIP(hipStreamSynchronize(ts));
  if (gpu_trace) {
    const double w0 = now_wall();
    parallel_for(n, [&](int i) {
      std::vector<uint32_t> &sc = cs[i]->sd_cells;
      const size_t nc = out_base[i + 1] - out_base[i];
      if (nc > sc.capacity()) {
        std::vector<uint32_t>().swap(sc);
        sc.reserve(nc);
        advise_huge(sc.data(), 4 * nc);
      }
      sc.assign(out_h + out_base[i], out_h + out_base[i + 1]);
      for (const auto &hs : host_slots[i]) std::copy(hs.second.begin(), hs.second.end(), sc.begin() + cs[i]->sd_off[hs.first]);
    });
    trace_ns += (int64_t)(1e9 * (now_wall() - w0));
  }
  tm.add_host(ST_DISC_LABELS, 1e-9 * (double)wait_ns.load());
  tm.add_host(ST_DISC_TRACE_CPU, 1e-9 * (double)trace_ns.load());
  tm.add_host(ST_DISC_GROUP_CPU, 1e-9 * (double)group_ns.load());
  tm.add_host(ST_DISC_GROUP, now_wall() - tp);
  if (*herr) throw std::logic_error("owner extremum not in surviving map");
}

}  // namespace mscx

extern "C" int mscx_disc_trace_stats(uint64_t out[5]) {
  return mscx::guarded([&] {
    MSCX_REQUIRE(out, MSCX_EINVAL, "null argument");
    out[0] = mscx::g_trace_gpu.load();
    out[1] = mscx::g_trace_host.load();
    out[2] = mscx::g_trace_dropped.load();
    out[3] = mscx::g_trace_wide.load();
    out[4] = mscx::g_trace_longest.load();
  });
}
