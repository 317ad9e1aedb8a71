// postMergeFillDiscs for many leaves at once: the critical-point pass and
// the saddle tracing run on host threads; the per-vertex / per-quad label
// pass (the bulk of the work: one owner lookup, redirects and an extremum
// map search per interior cell) runs on the GPU over every leaf in one
// launch, reading the owner images that are already resident in HBM.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "complex.hpp"
#include "runtime.hpp"
#include "tile.hpp"

namespace mscx {

static Scratch g_disc;   // host: codes + labels; device: descriptors, tables, labels
static Scratch g_stage;  // host staging of descriptors and tables

template <class F>
static void par_for(int n, F f) {
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

void fill_discs_device(mscx_tile *const *tiles, Complex *const *cs, int n) {
  if (n <= 0) return;
  ensure_init();
  Runtime &rt = runtime();
  std::lock_guard<std::mutex> lk(rt.mu);
  hipStream_t s = rt.stream;
  for (int i = 0; i < n; ++i) MSCX_REQUIRE(tiles[i]->worked, MSCX_ESTATE, "tile has not been worked");

  std::vector<Rect> rect(n), ext(n);
  std::vector<size_t> code_off(n + 1, 0), nv(n), nq(n);
  for (int i = 0; i < n; ++i) {
    const mscx_tile *t = tiles[i];
    rect[i] = Rect{t->rect[0], t->rect[1], t->rect[2], t->rect[3]};
    ext[i] = Rect{t->ext[0], t->ext[1], t->ext[2], t->ext[3]};
    code_off[i + 1] = code_off[i] + align_up((size_t)t->d.g.ch * t->d.pitch, 256);
    size_t nvx = (rect[i].x1 - rect[i].x0) / 2 + 1, nvy = (rect[i].y1 - rect[i].y0) / 2 + 1;
    nv[i] = nvx * nvy;
    nq[i] = (nvx - 1) * (nvy - 1);
  }

  // 1. gradient codes to pinned host memory (needed by the critical-point
  //    pass and the saddle tracing)
  std::vector<DiscPrep> dp(n);
  // sizes of the tables are known only after the critical-point pass, so
  // the codes live in their own region at the start of the host scratch
  size_t lab_total = 0;
  for (int i = 0; i < n; ++i) lab_total += align_up((nv[i] + nq[i]) * 4, 256);
  const size_t codes_bytes = code_off[n];
  char *host = (char *)g_disc.reserve_host(codes_bytes + lab_total);
  for (int i = 0; i < n; ++i)
    MSCX_HIP(hipMemcpyAsync(host + code_off[i], tiles[i]->d.codes, (size_t)tiles[i]->d.g.ch * tiles[i]->d.pitch,
                            hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipStreamSynchronize(s));
  std::vector<CodesView> cv(n);
  for (int i = 0; i < n; ++i)
    cv[i] = CodesView{(const uint8_t *)host + code_off[i], tiles[i]->d.pitch, ext[i].x0, ext[i].y0};

  // 2. critical-point pass (seeds, redirects, surviving maps)
  par_for(n, [&](int i) { disc_prepare(*cs[i], rect[i], cv[i], dp[i]); });

  // 3. device tables + label pass
  size_t tab_words = 0;
  for (int i = 0; i < n; ++i) tab_words += 2 * (dp[i].redirect.size() + dp[i].minmap.size() + dp[i].maxmap.size());
  const size_t o_desc = 0, o_pref = align_up(sizeof(DiscDesc) * n, 256);
  const size_t o_err = align_up(o_pref + 4 * (n + 1), 256);
  const size_t o_tab = o_err + 256;
  const size_t o_lab = align_up(o_tab + 4 * tab_words, 256);
  const size_t dev_bytes = o_lab + lab_total;
  char *dev = (char *)g_disc.reserve_dev(dev_bytes);
  char *stage = (char *)g_stage.reserve_host(o_lab);
  DiscDesc *hd = (DiscDesc *)(stage + o_desc);
  uint32_t *tab = (uint32_t *)(stage + o_tab);
  size_t tw = 0, lo = 0;
  std::vector<size_t> lab_off(n);
  for (int i = 0; i < n; ++i) {
    const mscx_tile *t = tiles[i];
    DiscDesc &D = hd[i];
    D.own = t->d.own;
    D.own_pitch = t->d.own_pitch;
    D.ox = t->ext[0];
    D.oy = t->ext[1];
    D.rx0 = rect[i].x0;
    D.ry0 = rect[i].y0;
    D.rx1 = rect[i].x1;
    D.ry1 = rect[i].y1;
    auto put = [&](const std::vector<std::pair<uint32_t, uint32_t>> &v, const uint32_t *&dptr, uint32_t &cnt) {
      dptr = (const uint32_t *)(dev + o_tab) + tw;
      cnt = (uint32_t)v.size();
      for (const auto &e : v) {
        tab[tw++] = e.first;
        tab[tw++] = e.second;
      }
    };
    put(dp[i].redirect, D.redir, D.nredir);
    put(dp[i].minmap, D.minmap, D.nmin);
    put(dp[i].maxmap, D.maxmap, D.nmax);
    lab_off[i] = lo;
    D.lab_v = (uint32_t *)(dev + o_lab + lo);
    D.lab_q = D.lab_v + nv[i];
    lo += align_up((nv[i] + nq[i]) * 4, 256);
    D.err = (uint32_t *)(dev + o_err);
  }
  int nblocks = plan_disc_blocks(hd, n, (int *)(stage + o_pref));
  *(uint32_t *)(stage + o_err) = 0;
  MSCX_HIP(hipMemcpyAsync(dev, stage, o_tab + 4 * tw, hipMemcpyHostToDevice, s));
  launch_disc_labels((const DiscDesc *)dev, (const int *)(dev + o_pref), n, nblocks, s);
  MSCX_HIP(hipGetLastError());
  char *labels = host + codes_bytes;
  MSCX_HIP(hipMemcpyAsync(labels, dev + o_lab, lo, hipMemcpyDeviceToHost, s));
  uint32_t *herr = (uint32_t *)(stage + o_err);
  MSCX_HIP(hipMemcpyAsync(herr, dev + o_err, 4, hipMemcpyDeviceToHost, s));
  MSCX_HIP(hipStreamSynchronize(s));
  if (*herr) throw std::logic_error("owner extremum not in surviving map");

  // 4. extremum discs (grouping) and saddle tracing, per leaf in parallel
  par_for(n, [&](int i) {
    const uint32_t *lv = (const uint32_t *)(labels + lab_off[i]);
    disc_group(*cs[i], rect[i], lv, lv + nv[i]);
    disc_paths(*cs[i], dp[i], rect[i], ext[i], cv[i]);
  });
}

}  // namespace mscx
