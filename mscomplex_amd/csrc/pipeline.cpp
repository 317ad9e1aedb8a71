// mscx_grid_run: the GridDataManager compute order (grid_datamanager.cpp:
// 596-662) on one GPU.  All leaf tiles go through the HIP stages in one
// batched launch per stage; the complex algebra runs on host threads, one
// task per independent merge of a tree level.
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "complex.hpp"
#include "runtime.hpp"
#include "tile.hpp"

namespace mscx {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class F>
static void parallel_for(int n, F f) {
  int nt = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
  if (n <= 1 || nt <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> errs(std::min(n, nt));
  for (int t = 0; t < (int)errs.size(); ++t)
    th.emplace_back([&, t] {
      try {
        for (int i = t; i < n; i += (int)errs.size()) f(i);
      } catch (...) {
        errs[t] = std::current_exception();
      }
    });
  for (auto &x : th) x.join();
  for (auto &e : errs)
    if (e) std::rethrow_exception(e);
}

struct LeafGeom {
  Rect rect, ext;
};

// createPieces_quadtree (grid_datamanager.cpp:53-89): split along y into
// 2^levels slabs; with lx/ly a 2-D tiling split along y while it has more
// levels left than x, else along x.  Children share the split line; each
// ext grows 2 cells (one vertex) past it.
static void create_pieces(Rect r, Rect e, int level, int lx, int ly, bool two_d, std::vector<LeafGeom> &out) {
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

}  // namespace mscx

using namespace mscx;

struct mscx_grid {
  std::vector<mscx_complex *> pieces;  // leaves, then internal nodes in merge order; root last
  std::vector<mscx_tile *> tiles;
  void *dev_input = nullptr;
  double t[5] = {0, 0, 0, 0, 0};
  ~mscx_grid() {
    for (mscx_complex *c : pieces) delete c;
    for (mscx_tile *t : tiles) mscx_tile_destroy(t);
    if (dev_input) mscx_dev_free(dev_input);
  }
};

extern "C" int mscx_grid_run(const float *data, int32_t sx, int32_t sy, int32_t levels, int32_t lx, int32_t ly,
                             double threshold, uint32_t flags, mscx_grid **out) {
  return guarded([&] {
    MSCX_REQUIRE(data && out, MSCX_EINVAL, "null argument");
    MSCX_REQUIRE(sx >= 2 && sy >= 2 && sx <= 16384 && sy <= 16384, MSCX_EINVAL, "grid sides must be in [2, 16384]");
    MSCX_REQUIRE(levels >= 0 && lx >= 0 && ly >= 0 && levels <= 12 && lx + ly <= 12, MSCX_EINVAL, "bad tiling");
    MSCX_REQUIRE(threshold >= 0.0 && !std::isnan(threshold), MSCX_EINVAL, "threshold must be >= 0");
    ensure_init();
    std::unique_ptr<mscx_grid> G(new mscx_grid);
    const double t0 = now_s();

    const float *dev = data;
    if (!(flags & MSCX_RUN_DEVICE_INPUT)) {
      size_t n = (size_t)sx * sy;
      for (size_t i = 0; i < n; ++i) MSCX_REQUIRE(!std::isnan(data[i]), MSCX_EINVAL, "input contains NaN");
      int rc = mscx_dev_alloc(n * 4, &G->dev_input);
      if (rc) throw Error(rc, mscx_last_error());
      rc = mscx_dev_upload(G->dev_input, data, n * 4);
      if (rc) throw Error(rc, mscx_last_error());
      dev = (const float *)G->dev_input;
    }

    bool two_d = lx > 0 || ly > 0;
    int L = two_d ? lx + ly : levels;
    std::vector<LeafGeom> leaves;
    Rect whole{0, 0, 2 * (sx - 1), 2 * (sy - 1)};
    create_pieces(whole, whole, L, lx, ly, two_d, leaves);
    const int nleaf = (int)leaves.size();
    for (const LeafGeom &lg : leaves) {
      int32_t r[4] = {lg.rect.x0, lg.rect.y0, lg.rect.x1, lg.rect.y1};
      int32_t e[4] = {lg.ext.x0, lg.ext.y0, lg.ext.x1, lg.ext.y1};
      mscx_tile *t = nullptr;
      int rc = mscx_tile_create(r, e, &t);
      if (rc) throw Error(rc, mscx_last_error());
      G->tiles.push_back(t);
      const float *win = dev + (size_t)(e[1] / 2) * sx + e[0] / 2;
      rc = mscx_tile_bind_device(t, win, sx);
      if (rc) throw Error(rc, mscx_last_error());
    }

    // computeSubdomainMsgraphs: GPU stages for every leaf, then the complexes
    run_tiles(G->tiles.data(), nleaf);
    std::vector<std::vector<uint32_t>> crit(nleaf), conn(nleaf), pr(nleaf);
    std::vector<std::vector<float>> fn(nleaf);
    for (int i = 0; i < nleaf; ++i) tile_critical(G->tiles[i], crit[i], fn[i], conn[i], pr[i]);
    G->pieces.assign(nleaf, nullptr);
    for (int i = 0; i < nleaf; ++i) G->pieces[i] = new mscx_complex;
    parallel_for(nleaf, [&](int i) {
      complex_build(G->pieces[i]->c, leaves[i].rect, leaves[i].ext, crit[i], fn[i], conn[i], pr[i]);
      std::vector<uint32_t>().swap(crit[i]);
      std::vector<uint32_t>().swap(conn[i]);
      std::vector<uint32_t>().swap(pr[i]);
      std::vector<float>().swap(fn[i]);
    });
    const double t1 = now_s();
    G->t[0] = t1 - t0;

    // mergePiecesUp (grid_datamanager.cpp:258-306): parent of (j, j+1) is
    // nleaf + j/2; one level at a time, merges of a level in parallel.
    G->pieces.resize(2 * nleaf - 1, nullptr);
    for (int p = nleaf; p < 2 * nleaf - 1; ++p) G->pieces[p] = new mscx_complex;
    for (int lo = 0, width = nleaf; width > 1; lo += width, width /= 2) {
      parallel_for(width / 2, [&](int k) {
        int j = lo + 2 * k;
        merge_up(G->pieces[j]->c, G->pieces[j + 1]->c, G->pieces[nleaf + j / 2]->c);
      });
    }
    const double t2 = now_s();
    G->t[1] = t2 - t1;

    if (threshold > 0.0) {
      std::vector<std::pair<uint32_t, uint32_t>> canc;
      simplify(G->pieces.back()->c, canc, threshold);
      un_simplify(G->pieces.back()->c, canc);
    }
    const double t3 = now_s();
    G->t[2] = t3 - t2;

    // mergePiecesDown + finalMergeDownPiecesInRange (:308-395): parents
    // before children, one tree level at a time.
    if (nleaf > 1) {
      int total = 2 * nleaf - 1;
      // internal node p has children 2(p - nleaf), 2(p - nleaf) + 1; walk
      // levels from the root down
      std::vector<int> level{total - 1};
      while (!level.empty()) {
        parallel_for((int)level.size(), [&](int k) {
          int p = level[k];
          int c = 2 * (p - nleaf);
          merge_down(G->pieces[p]->c, G->pieces[c]->c, G->pieces[c + 1]->c);
        });
        std::vector<int> next;
        for (int p : level) {
          int c = 2 * (p - nleaf);
          if (c >= nleaf) next.push_back(c);
          if (c + 1 >= nleaf) next.push_back(c + 1);
        }
        level.swap(next);
      }
    }
    const double t4 = now_s();
    G->t[3] = t4 - t3;

    // collectSubdomainManifolds / postMergeFillDiscs (:397-490)
    if (!(flags & MSCX_RUN_NO_DISCS)) {
      std::vector<Complex *> cs(nleaf);
      for (int i = 0; i < nleaf; ++i) cs[i] = &G->pieces[i]->c;
      fill_discs_device(G->tiles.data(), cs.data(), nleaf);
    }
    G->t[4] = now_s() - t4;
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
    *out = g->pieces[i];
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
