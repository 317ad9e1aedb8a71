// mscx_grid_run / mscx_batch_run: the GridDataManager compute order
// (grid_datamanager.cpp:596-662) for one grid or for many independent grids
// at once.  Every leaf tile of every grid goes through each HIP stage in one
// launch; the complex algebra runs on host threads (merges of one tree level
// in parallel for a single grid, whole grids in parallel for a batch); the
// manifold label pass of every leaf is again one launch.  The stages are
// also driven by the serving loop (server.cpp).
#include "pipeline.hpp"
#include "dmerge.hpp"

#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace mscx {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// createPieces_quadtree (grid_datamanager.cpp:53-89): split along y into
// 2^levels slabs; with lx/ly a 2-D tiling split along y while it has more
// levels left than x, else along x.  Children share the split line; each
// ext grows 2 cells (one vertex) past it.
void create_pieces(Rect r, Rect e, int level, int lx, int ly, bool two_d, std::vector<LeafGeom> &out) {
  if (level == 0) {
    out.push_back({r, e});
    return;
  }
  int dim = 1;
  if (two_d) dim = (ly >= lx && ly > 0) ? 1 : 0;
  int s = dim ? (r.y1 - r.y0) : (r.x1 - r.x0);
  Rect r1 = r, r2 = r, e1 = e, e2 = e;
  if (dim) {
    r1.y1 = r.y1 - 2 * (s / 4);
    r2.y0 = r1.y1;
    e1.y1 = r1.y1 + 2;
    e2.y0 = r2.y0 - 2;
  } else {
    r1.x1 = r.x1 - 2 * (s / 4);
    r2.x0 = r1.x1;
    e1.x1 = r1.x1 + 2;
    e2.x0 = r2.x0 - 2;
  }
  int nlx = lx, nly = ly;
  if (two_d) (dim ? nly : nlx)--;
  create_pieces(r1, e1, level - 1, nlx, nly, two_d, out);
  create_pieces(r2, e2, level - 1, nlx, nly, two_d, out);
}

// Leaf rects of a tiling.  The halo rule grows a child's ext one vertex past
// its split line even when the split falls on the grid's edge (a slab
// thinner than four cells splits off an empty piece), so a tiling finer
// than the grid can give ext rects that leave it; the reference would read
// past its input buffer for those (readDataAndInit copies the ext window),
// so they are rejected here, before any data is read.
static void leaves_of(int32_t sx, int32_t sy, int32_t levels, int32_t lx, int32_t ly, std::vector<LeafGeom> &out) {
  const bool two_d = lx > 0 || ly > 0;
  const Rect whole{0, 0, 2 * (sx - 1), 2 * (sy - 1)};
  create_pieces(whole, whole, two_d ? lx + ly : levels, lx, ly, two_d, out);
  for (const LeafGeom &lg : out)
    MSCX_REQUIRE(lg.ext.x0 >= 0 && lg.ext.y0 >= 0 && lg.ext.x1 <= whole.x1 && lg.ext.y1 <= whole.y1, MSCX_EINVAL,
                 "tiling is finer than the grid: a piece's ext rect leaves the grid");
}

// Where the merge tree of a plain or file-driven run goes when the caller
// does not say: HBM for grids of at least MSCX_DEVICE_MERGE_MIN_VERTICES
// vertices (environment).
uint32_t merge_policy(int32_t sx, int32_t sy, uint32_t flags) {
  if (flags & (MSCX_RUN_DEVICE_MERGE | MSCX_RUN_HOST_MERGE)) return flags;
  int64_t min_v = (int64_t)1 << 25;
  if (const char *e = getenv("MSCX_DEVICE_MERGE_MIN_VERTICES")) min_v = atoll(e);
  return (int64_t)sx * sy >= min_v ? (flags | MSCX_RUN_DEVICE_MERGE) : flags;
}

static void check_grid_args(int32_t sx, int32_t sy, int32_t levels, int32_t lx, int32_t ly) {
  MSCX_REQUIRE(sx >= 2 && sy >= 2 && sx <= 16384 && sy <= 16384, MSCX_EINVAL, "grid sides must be in [2, 16384]");
  MSCX_REQUIRE(levels >= 0 && lx >= 0 && ly >= 0 && levels <= 12 && lx + ly <= 12, MSCX_EINVAL, "bad tiling");
  std::vector<LeafGeom> leaves;
  leaves_of(sx, sy, levels, lx, ly, leaves);
}

void prepare_grid(mscx_grid *G, const float *data, int32_t sx, int32_t sy, int32_t levels, int32_t lx, int32_t ly,
                  double threshold, uint32_t flags) {
  MSCX_REQUIRE(data, MSCX_EINVAL, "null input");
  check_grid_args(sx, sy, levels, lx, ly);
  MSCX_REQUIRE(threshold >= 0.0 && !std::isnan(threshold), MSCX_EINVAL, "threshold must be >= 0");
  G->sx = sx;
  G->sy = sy;
  G->threshold = threshold;
  G->flags = flags;
  G->ooc = (flags & MSCX_RUN_OUT_OF_CORE) != 0;
  G->dmerge = (flags & MSCX_RUN_DEVICE_MERGE) && !(flags & MSCX_RUN_HOST_MERGE);
  // MSCX_WORK_LIMIT in the environment bounds the simplification's work in
  // plain and batched runs as the server's work limit does (0 or unset: none)
  if (const char *wl = getenv("MSCX_WORK_LIMIT")) G->work_limit = std::max(0.0, atof(wl));
  const float *dev = data;
  if (!(flags & MSCX_RUN_DEVICE_INPUT)) {
    size_t n = (size_t)sx * sy;
    for (size_t i = 0; i < n; ++i) MSCX_REQUIRE(!std::isnan(data[i]), MSCX_EINVAL, "input contains NaN");
    G->dev_input = device_pool().alloc(n * 4);
    MSCX_HIP(hipMemcpy(G->dev_input, data, n * 4, hipMemcpyHostToDevice));
    dev = (const float *)G->dev_input;
  }
  leaves_of(sx, sy, levels, lx, ly, G->leaves);
  for (const LeafGeom &lg : G->leaves) {
    int32_t r[4] = {lg.rect.x0, lg.rect.y0, lg.rect.x1, lg.rect.y1};
    int32_t e[4] = {lg.ext.x0, lg.ext.y0, lg.ext.x1, lg.ext.y1};
    mscx_tile *t = nullptr;
    int rc = mscx_tile_create(r, e, &t);
    if (rc) throw Error(rc, mscx_last_error());
    G->tiles.push_back(t);
    rc = mscx_tile_bind_device(t, dev + (size_t)(e[1] / 2) * sx + e[0] / 2, sx);
    if (rc) throw Error(rc, mscx_last_error());
  }
}

size_t grid_device_bytes(int32_t sx, int32_t sy, int32_t levels, int32_t lx, int32_t ly, bool with_input) {
  check_grid_args(sx, sy, levels, lx, ly);
  std::vector<LeafGeom> leaves;
  leaves_of(sx, sy, levels, lx, ly, leaves);
  size_t b = with_input ? (size_t)sx * sy * 4 : 0;
  for (const LeafGeom &lg : leaves) {
    int32_t r[4] = {lg.rect.x0, lg.rect.y0, lg.rect.x1, lg.rect.y1};
    int32_t e[4] = {lg.ext.x0, lg.ext.y0, lg.ext.x1, lg.ext.y1};
    b += tile_fixed_bytes(r, e);
    const size_t cells = (size_t)(e[2] - e[0] + 1) * (e[3] - e[1] + 1);
    b += CritLayout(cells / 4).total;
  }
  return b;
}

// Host leaf complexes (writeout_connectivity) of the grids from one host
// copy of the tiles' critical block.
// Host staging for the tiles' critical block: one per concurrent caller
// (a batch's leaves are built from it while its trees run).
struct CritStage {
  Scratch s;
};
static std::mutex g_crit_mu;
static std::vector<std::unique_ptr<CritStage>> g_crit_free;
struct CritLease {
  std::unique_ptr<CritStage> st;
  CritLease() {
    std::lock_guard<std::mutex> lk(g_crit_mu);
    if (g_crit_free.empty()) {
      st.reset(new CritStage);
    } else {
      st = std::move(g_crit_free.back());
      g_crit_free.pop_back();
    }
  }
  ~CritLease() {
    std::lock_guard<std::mutex> lk(g_crit_mu);
    g_crit_free.push_back(std::move(st));
  }
};

// Every tile's critical data lives in one block: one copy to host.
static const char *copy_crit_block(const std::vector<mscx_grid *> &grids, CritLease &lease) {
  const mscx_tile *t0 = grids[0]->tiles[0];
  const size_t bytes = t0->crit_block_bytes;
  char *host = (char *)lease.st->s.reserve_host(bytes);
  MSCX_HIP(hipMemcpyAsync(host, t0->crit_block.get(), bytes, hipMemcpyDeviceToHost, runtime().stream));
  MSCX_HIP(hipStreamSynchronize(runtime().stream));
  return host;
}

static void alloc_host_leaves(mscx_grid *G) {
  G->pieces.assign(G->leaves.size(), nullptr);
  for (size_t i = 0; i < G->leaves.size(); ++i) G->pieces[i] = new mscx_complex;
}

// Leaf i's complex (writeout_connectivity) from the host copy of the block.
static void build_host_leaf(mscx_grid *G, int i, const char *host) {
  const mscx_tile *t = G->tiles[i];
  const char *b = host + t->crit_off;
  const CritLayout L(t->K);
  complex_build(G->pieces[i]->c, G->leaves[i].rect, G->leaves[i].ext, t->K, (const uint32_t *)b,
                (const float *)(b + L.fn), (const uint32_t *)(b + L.conn), (const uint32_t *)(b + L.pair));
}

static void build_host_leaves(std::vector<mscx_grid *> &grids) {
  const int ng = (int)grids.size();
  if (!ng) return;
  CritLease lease;
  const char *host = copy_crit_block(grids, lease);
  std::vector<std::pair<int, int>> leafs;  // (grid, leaf), largest first
  for (int gi = 0; gi < ng; ++gi) {
    alloc_host_leaves(grids[gi]);
    for (size_t i = 0; i < grids[gi]->leaves.size(); ++i) leafs.emplace_back(gi, (int)i);
  }
  std::stable_sort(leafs.begin(), leafs.end(), [&](const std::pair<int, int> &a, const std::pair<int, int> &b) {
    return grids[a.first]->tiles[a.second]->K > grids[b.first]->tiles[b.second]->K;
  });
  parallel_for((int)leafs.size(), [&](int k) { build_host_leaf(grids[leafs[k].first], leafs[k].second, host); });
}

// GPU stages of every leaf of every grid and the leaf complexes.  With
// `defer`, the host leaves are allocated but left for the caller to build
// from the returned host copy of the critical block (held by the lease).
static const char *grids_compute_impl(std::vector<mscx_grid *> &grids, CritLease *defer) {
  const int ng = (int)grids.size();
  if (!ng) return nullptr;
  const double t0 = now_s();
  std::vector<mscx_tile *> tiles;
  for (mscx_grid *G : grids) tiles.insert(tiles.end(), G->tiles.begin(), G->tiles.end());
  run_tiles(tiles.data(), (int)tiles.size());
  std::vector<mscx_grid *> host, dev;
  for (mscx_grid *G : grids) (G->dmerge ? dev : host).push_back(G);
  const char *staged = nullptr;
  if (!defer) {
    build_host_leaves(host);
  } else if (!host.empty()) {
    staged = copy_crit_block(grids, *defer);
    for (mscx_grid *G : host) alloc_host_leaves(G);
  }
  // device merge tree: the leaves are built in HBM from the critical block
  std::vector<std::pair<int, int>> leafs;
  for (int gi = 0; gi < (int)dev.size(); ++gi) {
    mscx_grid *G = dev[gi];
    const size_t nl = G->leaves.size();
    G->pieces.assign(2 * nl - 1, nullptr);
    G->dp.clear();
    G->dp.resize(2 * nl - 1);
    for (size_t i = 0; i < nl; ++i) {
      G->dp[i].reset(new DComplex);
      leafs.emplace_back(gi, (int)i);
    }
  }
  parallel_for((int)leafs.size(), [&](int k) {
    mscx_grid *G = dev[leafs[k].first];
    const int i = leafs[k].second;
    StreamLease sl;
    dcomplex_from_tile(G->tiles[i], G->leaves[i].rect, G->leaves[i].ext, *G->dp[i], sl.s);
  });
  const double t1 = now_s();
  for (mscx_grid *G : grids) G->t[0] += t1 - t0;
  return staged;
}

void grids_compute(std::vector<mscx_grid *> &grids) { grids_compute_impl(grids, nullptr); }

// The device tree could not be followed: the tree runs on the host from
// host leaves built from the tiles' critical block.
static void device_tree_fallback(mscx_grid *G, const DFallback &why) {
  g_dfallbacks++;
  if (getenv("MSCX_DEBUG_MERGE")) fprintf(stderr, "[mscx] device merge tree falls back to the host: %s\n", why.why.c_str());
  for (mscx_complex *&c : G->pieces) delete c, c = nullptr;
  G->dp.clear();
  G->dmerge = false;
  std::vector<mscx_grid *> one{G};
  build_host_leaves(one);
}

// A failing device merge is redone on the host path, which raises the
// error the reference meets first (device runtime failures excepted).
template <class F>
static void device_merge_step(F f) {
  try {
    f();
  } catch (const Error &e) {
    if (e.code == MSCX_EDEVICE) throw;
    throw DFallback(e.what());
  } catch (const std::logic_error &e) {
    throw DFallback(e.what());
  }
}

static void grid_merge_up_dev(mscx_grid *G) {
  const int nleaf = (int)G->leaves.size();
  const double t1 = now_s();
  for (int lo = 0, width = nleaf; width > 1; lo += width, width /= 2)
    parallel_for(width / 2, [&](int k) {
      const int j = lo + 2 * k;
      StreamLease sl;
      std::unique_ptr<DComplex> o(new DComplex);
      device_merge_step([&] { dmerge_up(*G->dp[j], *G->dp[j + 1], *o, sl.s); });
      G->dp[nleaf + j / 2] = std::move(o);
    });
  G->t[1] += now_s() - t1;
}

static void grid_simplify_down_dev(mscx_grid *G) {
  const int nleaf = (int)G->leaves.size();
  const int root = 2 * nleaf - 2;
  const double t2 = now_s();
  if (G->threshold > 0.0) {
    // the simplification is one sequential chain of heap pops: the root
    // comes to the host and goes back to HBM for the merge down
    StreamLease sl;
    G->pieces[root] = new mscx_complex;
    Complex &R = G->pieces[root]->c;
    dcomplex_download(*G->dp[root], R, sl.s);
    std::vector<std::pair<uint32_t, uint32_t>> canc;
    simplify(R, canc, G->threshold, G->work_limit);
    un_simplify(R, canc, G->work_limit);
    if (nleaf > 1) {
      std::unique_ptr<DComplex> up(new DComplex);
      dcomplex_upload(R, *up, sl.s);
      G->dp[root] = std::move(up);
      // merge_down uncancels the root's line pairs in HBM: the host copy is
      // made again when asked for
      delete G->pieces[root];
      G->pieces[root] = nullptr;
    } else {
      G->dp[root].reset();  // the host copy is the piece
    }
  }
  const double t3 = now_s();
  G->t[2] += t3 - t2;
  if (nleaf > 1) {
    std::vector<int> level{root};
    while (!level.empty()) {
      parallel_for((int)level.size(), [&](int k) {
        const int p = level[k];
        const int c = 2 * (p - nleaf);
        StreamLease sl;
        device_merge_step([&] { dmerge_down(*G->dp[p], *G->dp[c], *G->dp[c + 1], sl.s); });
      });
      std::vector<int> next;
      for (int p : level) {
        const int c = 2 * (p - nleaf);
        if (c >= nleaf) next.push_back(c);
        if (c + 1 >= nleaf) next.push_back(c + 1);
      }
      if (!next.empty() && 2 * (next[0] - nleaf) >= nleaf) std::sort(next.rbegin(), next.rend());
      else std::sort(next.begin(), next.end());
      level.swap(next);
    }
  }
  // the manifold pass works on host copies of the leaves
  if (!(G->flags & MSCX_RUN_NO_DISCS)) {
    const double tc = now_s();
    const int T = host_threads();
    parallel_for(nleaf, [&](int i) {
      if (G->pieces[i]) return;
      // the leaves are copied side by side: each with its share of threads
      ThreadCap cap(std::max(1, T / nleaf));
      StreamLease sl;
      std::unique_ptr<mscx_complex> c(new mscx_complex);
      dcomplex_download(*G->dp[i], c->c, sl.s);
      G->pieces[i] = c.release();
    });
    stage_timer().add_host(ST_LEAF_COPIES, now_s() - tc);
  }
  G->t[3] += now_s() - t3;
}

static void init_spill(mscx_grid *G);

// Host threads one merge of a tree level may use inside: the level's share
// of the caller's threads when the level's merges run side by side, else 3
// (a batch runs whole grids side by side).
static int merge_threads(bool inner_parallel, int threads, int merges) {
  return inner_parallel ? std::max(1, threads / std::max(1, merges)) : 3;
}

// mergePiecesUp (grid_datamanager.cpp:258-306): parent of (j, j+1) is nleaf + j/2
void grid_merge_up(mscx_grid *G, bool inner_parallel) {
  if (G->dmerge) {
    try {
      grid_merge_up_dev(G);
      return;
    } catch (const DFallback &e) {
      device_tree_fallback(G, e);
    }
  }
  const int nleaf = (int)G->leaves.size();
  const int nt = inner_parallel ? 0 : 1;
  const double t1 = now_s();
  G->pieces.resize(2 * nleaf - 1, nullptr);
  if (G->ooc) init_spill(G);
  for (int p = nleaf; p < 2 * nleaf - 1; ++p) G->pieces[p] = new mscx_complex;
  const int T = host_threads();
  for (int lo = 0, width = nleaf; width > 1; lo += width, width /= 2)
    parallel_for(
        width / 2,
        [&](int k) {
          int j = lo + 2 * k;
          ThreadCap cap(merge_threads(inner_parallel, T, width / 2));
          merge_up(restore_piece(G, j)->c, restore_piece(G, j + 1)->c, G->pieces[nleaf + j / 2]->c);
          // out of core: the children wait on the side until merge down
          spill_piece(G, j);
          spill_piece(G, j + 1);
        },
        nt);
  G->t[1] += now_s() - t1;
}

// simplify_un_simplify at the root, then mergePiecesDown +
// finalMergeDownPiecesInRange (:308-395): parents before children, one tree
// level at a time.  Each level lists its parents in the reference's
// single-threaded order (mergePiecesDown walks an internal level from the
// highest piece index down, finalMergeDownPiecesInRange the leaves'
// parents upwards), so that of several failing merges the one reported is
// the one the reference meets first.
void grid_simplify_down(mscx_grid *G, bool inner_parallel) {
  if (G->dmerge) {
    try {
      grid_simplify_down_dev(G);
      return;
    } catch (const DFallback &e) {
      // the whole tree again, on the host
      G->t[1] = G->t[2] = G->t[3] = 0;
      device_tree_fallback(G, e);
      grid_merge_up(G, inner_parallel);
    }
  }
  const int nleaf = (int)G->leaves.size();
  const int nt = inner_parallel ? 0 : 1;
  const double t2 = now_s();
  if (G->threshold > 0.0) {
    std::vector<std::pair<uint32_t, uint32_t>> canc;
    simplify(G->pieces.back()->c, canc, G->threshold, G->work_limit);
    un_simplify(G->pieces.back()->c, canc, G->work_limit);
  }
  const double t3 = now_s();
  G->t[2] += t3 - t2;
  if (nleaf > 1) {
    const int T = host_threads();
    std::vector<int> level{2 * nleaf - 2};
    while (!level.empty()) {
      parallel_for(
          (int)level.size(),
          [&](int k) {
            int p = level[k];
            int c = 2 * (p - nleaf);
            ThreadCap cap(merge_threads(inner_parallel, T, (int)level.size()));
            merge_down(restore_piece(G, p)->c, restore_piece(G, c)->c, restore_piece(G, c + 1)->c);
            spill_piece(G, p);  // out of core: the parent is final
          },
          nt);
      std::vector<int> next;
      for (int p : level) {
        int c = 2 * (p - nleaf);
        if (c >= nleaf) next.push_back(c);
        if (c + 1 >= nleaf) next.push_back(c + 1);
      }
      // an internal level whose children are internal runs downwards
      if (!next.empty() && 2 * (next[0] - nleaf) >= nleaf) std::sort(next.rbegin(), next.rend());
      else std::sort(next.begin(), next.end());
      level.swap(next);
    }
  }
  G->t[3] += now_s() - t3;
}

// collectSubdomainManifolds / postMergeFillDiscs (:397-490) for every leaf
// of every grid that asked for discs
void grids_discs(std::vector<mscx_grid *> &grids, CodesStage *pre) {
  const double t4 = now_s();
  std::vector<mscx_tile *> dt;
  std::vector<Complex *> dc;
  for (mscx_grid *G : grids) {
    if (G->flags & MSCX_RUN_NO_DISCS) continue;
    for (size_t i = 0; i < G->leaves.size(); ++i) {
      dt.push_back(G->tiles[i]);
      dc.push_back(&restore_piece(G, (int)i)->c);
    }
  }
  fill_discs_device(dt.data(), dc.data(), (int)dt.size(), pre);
  for (mscx_grid *G : grids)
    if (G->ooc && G->leaves.size() > 1)
      parallel_for((int)G->leaves.size(), [&](int i) { spill_piece(G, i); });
  const double t5 = now_s();
  for (mscx_grid *G : grids) G->t[4] += t5 - t4;
}

// ---------------------------------------------------------------- out of core

static std::atomic<uint64_t> g_spill_seq{0};

static std::string spill_path(const mscx_grid *G, int i) { return G->spill_dir + "/dp_msgraph_" + std::to_string(i); }

// Sizes the spill store and, with MSCX_SPILL_DIR, makes this grid's
// directory of archive files there.
static void init_spill(mscx_grid *G) {
  const size_t n = 2 * G->leaves.size() - 1;
  G->spilled.clear();
  G->spilled.resize(n);
  G->spilled_file.assign(n, 0);
  const char *dir = getenv("MSCX_SPILL_DIR");
  if (dir && *dir) {
    G->spill_dir = std::string(dir) + "/mscx_spill_" + std::to_string((long)getpid()) + "_" +
                   std::to_string(g_spill_seq.fetch_add(1));
    MSCX_REQUIRE(mkdir(G->spill_dir.c_str(), 0700) == 0, MSCX_EIO, "cannot create spill directory " + G->spill_dir);
  }
}

void spill_piece(mscx_grid *G, int i) {
  if (!G->ooc || i < 0 || (size_t)i >= G->spilled.size() || !G->pieces[i]) return;
  if (!G->spill_dir.empty()) {
    save_complex(G->pieces[i]->c, spill_path(G, i));
    G->spilled_file[i] = 1;
  } else {
    std::unique_ptr<ComplexSpill> sp(new ComplexSpill);
    spill_complex(G->pieces[i]->c, *sp);
    G->spilled[i] = std::move(sp);
  }
  delete G->pieces[i];
  G->pieces[i] = nullptr;
}

mscx_complex *restore_piece(mscx_grid *G, int i) {
  if (G->pieces[i]) return G->pieces[i];
  if (G->dmerge && (size_t)i < G->dp.size() && G->dp[i] &&
      !(G->ooc && (size_t)i < G->spilled.size() && (G->spilled[i] || G->spilled_file[i]))) {
    // device merge tree: a host copy of the piece's complex
    StreamLease sl;
    std::unique_ptr<mscx_complex> c(new mscx_complex);
    dcomplex_download(*G->dp[i], c->c, sl.s);
    G->pieces[i] = c.release();
    return G->pieces[i];
  }
  MSCX_REQUIRE(G->ooc && (size_t)i < G->spilled.size(), MSCX_ESTATE, "piece is neither live nor spilled");
  std::unique_ptr<mscx_complex> c(new mscx_complex);
  if (G->spilled_file[i]) {
    load_complex(spill_path(G, i), c->c);
    unlink(spill_path(G, i).c_str());
    G->spilled_file[i] = 0;
  } else {
    MSCX_REQUIRE(G->spilled[i], MSCX_ESTATE, "piece is neither live nor spilled");
    restore_complex(*G->spilled[i], c->c);
    G->spilled[i].reset();
  }
  G->pieces[i] = c.release();
  return G->pieces[i];
}

void copy_pieces(const mscx_grid *src, mscx_grid *dst) {
  for (mscx_complex *c : dst->pieces) delete c;
  dst->pieces.assign(src->pieces.size(), nullptr);
  dst->dp.clear();
  dst->dmerge = src->dmerge;
  if (src->dmerge) {
    // device-resident complexes: copied in HBM
    dst->dp.resize(src->dp.size());
    try {
      parallel_for((int)src->dp.size(), [&](int i) {
        if (!src->dp[i]) return;
        StreamLease sl;
        std::unique_ptr<DComplex> d(new DComplex);
        dcomplex_copy(*src->dp[i], *d, sl.s);
        dst->dp[i] = std::move(d);
      });
    } catch (...) {
      dst->dp.clear();
      throw;
    }
  }
  // the pieces are independent: copied on the host threads
  try {
    parallel_for((int)src->pieces.size(), [&](int i) {
      if (src->pieces[i]) dst->pieces[i] = new mscx_complex(*src->pieces[i]);
    });
  } catch (...) {
    for (mscx_complex *&c : dst->pieces) delete c, c = nullptr;
    dst->pieces.clear();
    throw;
  }
}

size_t device_tree_bytes(const mscx_grid *G) {
  size_t b = 0;
  for (const auto &d : G->dp)
    if (d) b += d->device_bytes();
  return b;
}

}  // namespace mscx

using namespace mscx;

mscx_grid::~mscx_grid() {
  // releasing a large grid's complexes (millions of small blocks) is spread
  // over the host threads; it completes before the destructor returns
  if (pieces.size() >= 8) {
    try {
      parallel_for((int)pieces.size(), [&](int i) {
        delete pieces[i];
        pieces[i] = nullptr;
      });
    } catch (...) {  // thread creation failed: release the rest here
      for (mscx_complex *&c : pieces) delete c, c = nullptr;
    }
  } else {
    for (mscx_complex *c : pieces) delete c;
  }
  if (owns_tiles)
    for (mscx_tile *t : tiles) mscx_tile_destroy(t);
  if (dev_input) mscx_dev_free(dev_input);
  if (!spill_dir.empty()) {
    for (size_t i = 0; i < spilled_file.size(); ++i)
      if (spilled_file[i]) unlink((spill_dir + "/dp_msgraph_" + std::to_string(i)).c_str());
    rmdir(spill_dir.c_str());
  }
}

struct mscx_batch {
  std::vector<mscx_grid *> grids;
  std::vector<int> status;
  std::vector<std::string> error;
  ~mscx_batch() {
    for (mscx_grid *g : grids) delete g;
  }
};

namespace mscx {

// The whole compute order for a set of prepared grids.  A failure of one
// grid in the host phases is recorded in status/error (when given) so the
// other grids complete; device failures fail the call.
static void run_grids(std::vector<mscx_grid *> &grids, std::vector<int> *status, std::vector<std::string> *error) {
  const int ng = (int)grids.size();
  if (!ng) return;
  // A batch builds each grid's host leaves in the task that then runs its
  // tree, so that a large leaf's build overlaps other grids' trees instead
  // of holding every tree back.
  CritLease lease;
  const char *staged = grids_compute_impl(grids, ng > 1 ? &lease : nullptr);
  // one grid: its codes come to the host while the host runs its merge tree
  std::shared_ptr<CodesStage> codes;
  if (ng == 1 && !(grids[0]->flags & MSCX_RUN_NO_DISCS))
    codes = stage_codes(grids[0]->tiles.data(), (int)grids[0]->leaves.size());
  std::vector<char> ok(ng, 1);
  auto tree = [&](int gi, bool inner) {
    try {
      mscx_grid *G = grids[gi];
      if (staged && !G->dmerge) {
        const double t0 = now_s();
        for (int i = 0; i < (int)G->leaves.size(); ++i) build_host_leaf(G, i, staged);
        G->t[0] += now_s() - t0;
      }
      grid_merge_up(grids[gi], inner);
      grid_simplify_down(grids[gi], inner);
    } catch (const std::exception &e) {
      if (!status) throw;
      ok[gi] = 0;
      int code = MSCX_ELOGIC;
      if (auto *er = dynamic_cast<const Error *>(&e)) code = er->code;
      (*status)[gi] = code;
      (*error)[gi] = e.what();
    }
  };
  if (ng == 1) {
    tree(0, true);
  } else {
    // largest grids first: the batch ends when its slowest tree does, so
    // the long ones start at once and the small ones fill the gaps
    std::vector<int> order(ng);
    for (int gi = 0; gi < ng; ++gi) order[gi] = gi;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return (int64_t)grids[a]->sx * grids[a]->sy > (int64_t)grids[b]->sx * grids[b]->sy;
    });
    parallel_for(ng, [&](int k) { tree(order[k], false); });
  }
  std::vector<mscx_grid *> done;
  for (int gi = 0; gi < ng; ++gi)
    if (ok[gi]) done.push_back(grids[gi]);
  grids_discs(done, codes.get());
}

}  // namespace mscx

extern "C" int mscx_grid_run(const float *data, int32_t sx, int32_t sy, int32_t levels, int32_t lx, int32_t ly,
                             double threshold, uint32_t flags, mscx_grid **out) {
  return guarded([&] {
    MSCX_REQUIRE(out && data, MSCX_EINVAL, "null argument");
    check_grid_args(sx, sy, levels, lx, ly);  // before the device is touched
    ensure_init();
    std::unique_ptr<mscx_grid> G(new mscx_grid);
    const double t0 = now_s();
    prepare_grid(G.get(), data, sx, sy, levels, lx, ly, threshold, merge_policy(sx, sy, flags));
    G->t[0] += now_s() - t0;
    std::vector<mscx_grid *> gs{G.get()};
    run_grids(gs, nullptr, nullptr);
    *out = G.release();
  });
}

extern "C" int mscx_batch_run(int32_t n, const float *const *data, const int32_t *sx, const int32_t *sy,
                              const int32_t *lx, const int32_t *ly, const double *threshold, uint32_t flags,
                              mscx_batch **out) {
  return guarded([&] {
    MSCX_REQUIRE(out && n >= 0 && (n == 0 || (data && sx && sy && lx && ly && threshold)), MSCX_EINVAL,
                 "null argument");
    ensure_init();
    std::unique_ptr<mscx_batch> B(new mscx_batch);
    B->status.assign(n, MSCX_OK);
    B->error.assign(n, std::string());
    std::vector<mscx_grid *> live;
    for (int i = 0; i < n; ++i) {
      mscx_grid *G = new mscx_grid;
      B->grids.push_back(G);
      try {
        prepare_grid(G, data[i], sx[i], sy[i], 0, lx[i], ly[i], threshold[i], flags);
        live.push_back(G);
      } catch (const Error &e) {
        B->status[i] = e.code;
        B->error[i] = e.what();
      }
    }
    std::vector<int> st(live.size(), MSCX_OK);
    std::vector<std::string> er(live.size());
    run_grids(live, &st, &er);
    for (size_t k = 0, i = 0; i < (size_t)n; ++i) {
      if (B->status[i] != MSCX_OK) continue;
      B->status[i] = st[k];
      B->error[i] = er[k];
      ++k;
    }
    *out = B.release();
  });
}

extern "C" int mscx_batch_size(const mscx_batch *b, int32_t *n) {
  return guarded([&] {
    MSCX_REQUIRE(b && n, MSCX_EINVAL, "null argument");
    *n = (int32_t)b->grids.size();
  });
}

extern "C" int mscx_batch_grid(const mscx_batch *b, int32_t i, const mscx_grid **out, int32_t *status) {
  return guarded([&] {
    MSCX_REQUIRE(b && out && i >= 0 && i < (int32_t)b->grids.size(), MSCX_EINVAL, "bad grid index");
    *out = b->grids[i];
    if (status) *status = b->status[i];
  });
}

extern "C" const char *mscx_batch_error(const mscx_batch *b, int32_t i) {
  if (!b || i < 0 || i >= (int32_t)b->grids.size()) return "";
  return b->error[i].c_str();
}

extern "C" int mscx_batch_destroy(mscx_batch *b) {
  return guarded([&] { delete b; });
}

// GridDataManager's constructor over a file (grid_datamanager.cpp:596-662):
// read, compute, merge up, simplify (threshold > 0), merge down, collect the
// manifolds and write each leaf's discs (collectManifold, :397-420).  The
// file is checked before the device is touched.
extern "C" int mscx_grid_run_file(const char *path, int32_t sx, int32_t sy, int32_t levels, int32_t lx, int32_t ly,
                                  double threshold, int32_t num_parallel, uint32_t flags, const char *disc_prefix,
                                  mscx_grid **out) {
  return guarded([&] {
    MSCX_REQUIRE(path && out, MSCX_EINVAL, "null argument");
    MSCX_REQUIRE(!(flags & MSCX_RUN_DEVICE_INPUT), MSCX_EINVAL, "a file run reads host input");
    MSCX_REQUIRE(!(disc_prefix && (flags & MSCX_RUN_NO_DISCS)), MSCX_EINVAL, "disc files need the discs");
    MSCX_REQUIRE(num_parallel >= 0 && num_parallel <= 1024, MSCX_EINVAL, "num_parallel must be in [0, 1024]");
    MSCX_REQUIRE(threshold >= 0.0 && !std::isnan(threshold), MSCX_EINVAL, "threshold must be >= 0");
    check_grid_args(sx, sy, levels, lx, ly);
    GridFile f;
    open_grid_file(path, sx, sy, f);
    ensure_init();
    ThreadCap cap(num_parallel);
    std::unique_ptr<mscx_grid> G(new mscx_grid);
    const double t0 = now_s();
    G->dev_input = device_pool().alloc(f.bytes);
    stream_grid_file(f, sx, (float *)G->dev_input);
    prepare_grid(G.get(), (const float *)G->dev_input, sx, sy, levels, lx, ly, threshold,
                 merge_policy(sx, sy, flags) | MSCX_RUN_DEVICE_INPUT);
    G->t[0] += now_s() - t0;
    std::vector<mscx_grid *> gs{G.get()};
    run_grids(gs, nullptr, nullptr);
    if (disc_prefix) {
      const double t5 = now_s();
      for (size_t i = 0; i < G->leaves.size(); ++i) {
        write_discs(restore_piece(G.get(), (int)i)->c, disc_prefix);
        if (G->leaves.size() > 1) spill_piece(G.get(), (int)i);
      }
      G->t[4] += now_s() - t5;
    }
    *out = G.release();
  });
}

extern "C" int mscx_grid_num_pieces(const mscx_grid *g, int32_t *n) {
  return guarded([&] {
    MSCX_REQUIRE(g && n, MSCX_EINVAL, "null argument");
    *n = (int32_t)g->pieces.size();
  });
}

extern "C" int mscx_grid_piece(const mscx_grid *g, int32_t i, const mscx_complex **out) {
  return guarded([&] {
    MSCX_REQUIRE(g && out && i >= 0 && i < (int32_t)g->pieces.size(), MSCX_EINVAL, "bad piece index");
    if (!g->pieces[i]) {
      // out of core: the piece comes back into memory and stays there
      mscx_grid *G = const_cast<mscx_grid *>(g);
      std::lock_guard<std::mutex> lk(G->spill_mu);
      restore_piece(G, i);
    }
    *out = g->pieces[i];
  });
}

extern "C" int mscx_grid_piece_num_cps(const mscx_grid *g, int32_t i, uint32_t *n) {
  return guarded([&] {
    MSCX_REQUIRE(g && n && i >= 0 && i < (int32_t)g->pieces.size(), MSCX_EINVAL, "bad piece index");
    if (g->pieces[i]) *n = g->pieces[i]->c.size();
    else if ((size_t)i < g->dp.size() && g->dp[i]) *n = g->dp[i]->n;
    else {
      mscx_grid *G = const_cast<mscx_grid *>(g);
      std::lock_guard<std::mutex> lk(G->spill_mu);
      *n = restore_piece(G, i)->c.size();
    }
  });
}

// FNV-1a over the dump words of every piece, in piece order.
extern "C" int mscx_grid_checksum(const mscx_grid *g, uint64_t *out) {
  return guarded([&] {
    MSCX_REQUIRE(g && out, MSCX_EINVAL, "null argument");
    mscx_grid *G = const_cast<mscx_grid *>(g);
    uint64_t h = 1469598103934665603ull;
    std::vector<uint32_t> w;
    for (size_t i = 0; i < G->pieces.size(); ++i) {
      const mscx_complex *c = G->pieces[i];
      if (!c) {
        std::lock_guard<std::mutex> lk(G->spill_mu);
        c = restore_piece(G, (int)i);
      }
      c->c.dump(w);
      const unsigned char *b = (const unsigned char *)w.data();
      for (size_t k = 0; k < w.size() * 4; ++k) h = (h ^ b[k]) * 1099511628211ull;
    }
    *out = h;
  });
}

extern "C" int mscx_grid_times(const mscx_grid *g, double t5[5]) {
  return guarded([&] {
    MSCX_REQUIRE(g && t5, MSCX_EINVAL, "null argument");
    memcpy(t5, g->t, sizeof(g->t));
  });
}

extern "C" int mscx_grid_destroy(mscx_grid *g) {
  return guarded([&] { delete g; });
}
