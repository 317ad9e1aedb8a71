// oracle/oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY.
// C entry points so the Python tests and the CPU-baseline leg of bench.py can
// drive the oracle through ctypes.  Arrays over a tile are laid out over the
// tile's ext rect, x fastest: index = (x - ext.x0) + (y - ext.y0) * (ext.x1 - ext.x0 + 1).
#include <cstring>
#include <stdexcept>
#include <string>

#include "oracle.hpp"

using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

// Selects a planted mistake (0 = none; the values are listed at g_plant in
// oracle.cpp).  Only the pin tests call it, to show their checks fail on
// each mistake.
void orc_set_plant(int v, int stop_after) {
  orc::g_plant = v;
  orc::g_stop_after = stop_after;
}

// GridDataset(rect, ext).init(data).work() for one tile.  crit receives the
// critical list (packed cellids, x | y<<16) when crit_cap suffices; *K gets
// its length either way.
int orc_tile_work(const float *data, const int *rect, const int *ext, int8_t *flags, uint32_t *pairs,
                  uint32_t *own, uint32_t *crit, size_t crit_cap, size_t *K) {
  try {
    GridDataset ds(rect_t(rect[0], rect[1], rect[2], rect[3]), rect_t(ext[0], ext[1], ext[2], ext[3]));
    ds.init(data);
    ds.work();
    int W = ext[2] - ext[0] + 1, H = ext[3] - ext[1] + 1;
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        cellid_t c(ext[0] + x, ext[1] + y);
        size_t o = (size_t)x + (size_t)y * W, i = ds.cidx(c);
        if (flags) flags[o] = ds.m_cell_flags[i];
        if (pairs) pairs[o] = ds.m_cell_pairs[i].packed();
        if (own) own[o] = ds.m_cell_own[i].packed();
      }
    *K = ds.m_critical_cells.size();
    if (crit && crit_cap >= *K)
      for (size_t k = 0; k < *K; ++k) crit[k] = ds.m_critical_cells[k].packed();
    return 0;
  } catch (std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// work() + writeout_connectivity() for one tile; writes the complex dump.
// Returns the dump length in uint32 words (the dump is written only when
// cap suffices), or -1 on error.
long orc_tile_complex(const float *data, const int *rect, const int *ext, uint32_t *buf, size_t cap) {
  try {
    GridDataset ds(rect_t(rect[0], rect[1], rect[2], rect[3]), rect_t(ext[0], ext[1], ext[2], ext[3]));
    GridMSComplex msc(ds.m_rect, ds.m_ext_rect);
    ds.init(data);
    ds.work();
    ds.writeout_connectivity(&msc);
    std::vector<uint32_t> d;
    msc.dump(d);
    if (buf && cap >= d.size()) memcpy(buf, d.data(), d.size() * 4);
    return (long)d.size();
  } catch (std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

struct OrcPipeline {
  DataManager *dm = nullptr;
  PipelineSnapshots snaps;
  int nleaves = 0;
};

// The whole GridDataManager compute order on a size_x x size_y grid (x
// fastest).  levels: slab levels (split along y); lx/ly > 0 selects a 2-D
// tiling of 2^lx x 2^ly tiles instead.  keep != 0 stores complex dumps of
// every piece after every phase.  nthreads > 1 runs the independent pieces
// of each phase on that many threads (the reference's --num-parallel).
void *orc_pipeline_run(const float *data, int sx, int sy, int levels, int lx, int ly, double thresh, int keep,
                       int nthreads) {
  OrcPipeline *p = new OrcPipeline;
  try {
    p->dm = new DataManager(data, sx, sy, levels, lx, ly, thresh, keep != 0);
    p->nleaves = 1 << p->dm->m_num_levels;
    p->dm->run(&p->snaps, nthreads);
    return p;
  } catch (std::exception &e) {
    g_err = e.what();
    delete p->dm;
    delete p;
    return nullptr;
  }
}

int orc_pipeline_npieces(void *h) { return (int)((OrcPipeline *)h)->dm->m_pieces.size(); }
int orc_pipeline_nleaves(void *h) { return ((OrcPipeline *)h)->nleaves; }

// stage: 0 after compute, 1 after merge-up, 2 after simplify, 3 after
// merge-down, 4 after manifold discs.
long orc_pipeline_dump(void *h, int stage, int piece, uint32_t *buf, size_t cap) {
  OrcPipeline *p = (OrcPipeline *)h;
  std::vector<std::vector<uint32_t>> *st[] = {&p->snaps.after_compute, &p->snaps.after_merge_up,
                                              &p->snaps.after_simplify, &p->snaps.after_merge_down,
                                              &p->snaps.after_discs};
  if (stage < 0 || stage > 4) return -1;
  std::vector<std::vector<uint32_t>> &v = *st[stage];
  if (piece < 0 || (size_t)piece >= v.size()) return 0;
  const std::vector<uint32_t> &d = v[piece];
  if (buf && cap >= d.size()) memcpy(buf, d.data(), d.size() * 4);
  return (long)d.size();
}

void orc_pipeline_rect(void *h, int piece, int *rect_ext8) {
  OrcPipeline *p = (OrcPipeline *)h;
  const Piece *dp = p->dm->m_pieces[piece];
  const GridMSComplex *m = dp->msgraph;
  const rect_t *rs[2] = {&m->m_rect, &m->m_ext_rect};
  for (int k = 0; k < 2; ++k) {
    rect_ext8[4 * k + 0] = rs[k]->bl[0];
    rect_ext8[4 * k + 1] = rs[k]->bl[1];
    rect_ext8[4 * k + 2] = rs[k]->tr[0];
    rect_ext8[4 * k + 3] = rs[k]->tr[1];
  }
}

void orc_pipeline_times(void *h, double *t5) {
  OrcPipeline *p = (OrcPipeline *)h;
  t5[0] = p->snaps.t_compute;
  t5[1] = p->snaps.t_merge_up;
  t5[2] = p->snaps.t_simplify;
  t5[3] = p->snaps.t_merge_down;
  t5[4] = p->snaps.t_manifolds;
}

void orc_pipeline_free(void *h) {
  OrcPipeline *p = (OrcPipeline *)h;
  if (!p) return;
  delete p->dm;
  delete p;
}

}  // extern "C"
