// The product's GridMSComplex algorithms (see complex.hpp).
#include "complex.hpp"
#include "heap.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <queue>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "runtime.hpp"
#include "tile.hpp"

namespace mscx {

// ------------------------------------------------------------------ CellMap

void CellMap::reserve(size_t n) {
  size_t want = 16;
  while (want < n * 2) want <<= 1;
  if (want <= slots_.size()) return;
  std::vector<Slot> old;
  old.swap(slots_);
  slots_.reserve(want);
  advise_huge(slots_.data(), want * sizeof(Slot));
  slots_.assign(want, Slot{UINT32_MAX, 0});
  mask_ = want - 1;
  n_ = 0;
  for (const Slot &e : old)
    if (e.key != UINT32_MAX) set(e.key, e.val);
}

void *CellMap::raw_table(size_t slots, size_t n) {
  if (slots > slots_.capacity()) {
    std::vector<Slot>().swap(slots_);
    slots_.reserve(slots);
    advise_huge(slots_.data(), slots * sizeof(Slot));
  }
  slots_.resize(slots);
  mask_ = slots - 1;
  n_ = n;
  return slots_.data();
}

void CellMap::grow() { reserve(slots_.empty() ? 16 : slots_.size()); }

size_t CellMap::slot(uint32_t key) const {
  size_t i = hash(key) & mask_;
  const Slot *t = slots_.data();
  while (t[i].key != UINT32_MAX && t[i].key != key) i = (i + 1) & mask_;
  return i;
}

uint32_t CellMap::find(uint32_t key) const {
  if (slots_.empty()) return UINT32_MAX;
  const Slot &e = slots_[slot(key)];
  return e.key == key ? e.val : UINT32_MAX;
}

uint32_t CellMap::find_or_insert(uint32_t key, uint32_t val) {
  if ((n_ + 1) * 2 > slots_.size()) grow();
  Slot &e = slots_[slot(key)];
  if (e.key == key) return e.val;
  e = Slot{key, val};
  ++n_;
  return UINT32_MAX;
}

void CellMap::set(uint32_t key, uint32_t val) {
  if ((n_ + 1) * 2 > slots_.size()) grow();
  Slot &e = slots_[slot(key)];
  if (e.key != key) {
    e.key = key;
    ++n_;
  }
  e.val = val;
}

uint32_t CellMap::get_or_zero(uint32_t key) {
  const uint32_t v = find_or_insert(key, 0);
  return v == UINT32_MAX ? 0 : v;
}

// ------------------------------------------------------------------ Complex

DiscSpan Complex::disc(uint32_t i, int j) const {
  if (i < asc_disc.size()) {
    const std::vector<uint32_t> &v = j == 0 ? asc_disc[i] : des_disc[i];
    if (!v.empty()) return DiscSpan{v.data(), v.size()};
  }
  const size_t sl = 2 * (size_t)i + j;
  if (sl + 1 < sd_off.size() && sd_off[sl + 1] > sd_off[sl])
    return DiscSpan{sd_cells.data() + sd_off[sl], sd_off[sl + 1] - sd_off[sl]};
  uint32_t d = cell_dim(cell[i]);
  if (!xd_off.empty() && i + 1 < xd_off.size() && ((j == 0 && d == 0) || (j == 1 && d == 2)))
    return DiscSpan{xd_cells.data() + xd_off[i], xd_off[i + 1] - xd_off[i]};
  return DiscSpan{nullptr, 0};
}

template <class V>
static void reserve_huge(V &v, size_t n) {
  if (n <= v.capacity()) return;
  v.reserve(n);
  advise_huge(v.data(), v.capacity() * sizeof(typename V::value_type));
}

void Complex::reserve_nodes(size_t n) {
  reserve_huge(cell, n);
  reserve_huge(fn, n);
  reserve_huge(pair, n);
  reserve_huge(flags, n);
  reserve_huge(asc, n);
  reserve_huge(des, n);
}

uint32_t Complex::replicate(const Complex &src, uint32_t j) {
  uint32_t c = src.cell[j];
  uint32_t n = size();
  if (idx.find_or_insert(c, n) != UINT32_MAX) throw std::logic_error("this cp is present in msc");
  cell.push_back(c);
  pair.push_back(UINT32_MAX);
  flags.push_back(src.flags[j]);
  asc.emplace_back();
  des.emplace_back();
  if (!asc_disc.empty()) {
    asc_disc.emplace_back();
    des_disc.emplace_back();
  }
  return n;
}

void Complex::dump(std::vector<uint32_t> &out) const {
  out.clear();
  out.push_back(size());
  out.push_back((uint32_t)rect.x0);
  out.push_back((uint32_t)rect.y0);
  out.push_back((uint32_t)rect.x1);
  out.push_back((uint32_t)rect.y1);
  out.push_back((uint32_t)ext.x0);
  out.push_back((uint32_t)ext.y0);
  out.push_back((uint32_t)ext.x1);
  out.push_back((uint32_t)ext.y1);
  for (uint32_t i = 0; i < size(); ++i) {
    uint32_t fb;
    float f = i < fn.size() ? fn[i] : 0.0f;
    memcpy(&fb, &f, 4);
    out.push_back(cell[i]);
    out.push_back(fb);
    out.push_back(pair[i]);
    out.push_back(flags[i]);
    out.push_back(asc[i].size());
    out.push_back(des[i].size());
    DiscSpan ad = disc(i, 0), dd = disc(i, 1);
    out.push_back((uint32_t)ad.n);
    out.push_back((uint32_t)dd.n);
    out.insert(out.end(), asc[i].begin(), asc[i].end());
    out.insert(out.end(), des[i].begin(), des[i].end());
    out.insert(out.end(), ad.p, ad.p + ad.n);
    out.insert(out.end(), dd.p, dd.p + dd.n);
  }
}

// ---------------------------------------------------------------- spill

// Node i's disc list j as held per node (loaded from an archive), or an
// empty list when the complex holds none per node.
static const std::vector<uint32_t> &node_disc(const Complex &m, uint32_t i, int j) {
  static const std::vector<uint32_t> none;
  if (i >= m.asc_disc.size()) return none;
  return j == 0 ? m.asc_disc[i] : m.des_disc[i];
}

// Nodes per record chunk of a spilled complex, and the size from which its
// chunks are written and read on several host threads.
static constexpr uint32_t kSpillChunk = 1u << 20;

void spill_complex(Complex &m, ComplexSpill &out) {
  const uint32_t n = m.size();
  out.rect = m.rect;
  out.ext = m.ext;
  out.n = n;
  const uint32_t nc = (n + kSpillChunk - 1) / kSpillChunk;
  out.begin.resize(nc + 1);
  for (uint32_t c = 0; c <= nc; ++c) out.begin[c] = std::min(n, c * kSpillChunk);
  out.words.assign(nc, {});
  parallel_for((int)nc, [&](int c) {
    std::vector<uint32_t> &w = out.words[c];
    size_t len = 0;
    for (uint32_t i = out.begin[c]; i < out.begin[c + 1]; ++i)
      len += 8 + m.asc[i].size() + m.des[i].size() + node_disc(m, i, 0).size() + node_disc(m, i, 1).size();
    w.reserve(len);
    for (uint32_t i = out.begin[c]; i < out.begin[c + 1]; ++i) {
      uint32_t fb;
      memcpy(&fb, &m.fn[i], 4);
      const uint32_t rec[8] = {m.cell[i], fb, m.pair[i], m.flags[i], m.asc[i].size(), m.des[i].size(),
                               (uint32_t)node_disc(m, i, 0).size(), (uint32_t)node_disc(m, i, 1).size()};
      w.insert(w.end(), rec, rec + 8);
      w.insert(w.end(), m.asc[i].begin(), m.asc[i].end());
      w.insert(w.end(), m.des[i].begin(), m.des[i].end());
      w.insert(w.end(), node_disc(m, i, 0).begin(), node_disc(m, i, 0).end());
      w.insert(w.end(), node_disc(m, i, 1).begin(), node_disc(m, i, 1).end());
    }
  });
  out.node_discs = !m.asc_disc.empty();
  out.idx = std::move(m.idx);
  out.xd_off = std::move(m.xd_off);
  out.xd_cells = std::move(m.xd_cells);
  out.sd_off = std::move(m.sd_off);
  out.sd_cells = std::move(m.sd_cells);
  m = Complex();
}

void restore_complex(ComplexSpill &in, Complex &m) {
  Complex c;
  const uint32_t n = in.n;
  c.rect = in.rect;
  c.ext = in.ext;
  c.reserve_nodes(Complex::with_headroom(n));
  c.cell.resize(n);
  c.fn.resize(n);
  c.pair.resize(n);
  c.flags.resize(n);
  c.asc.resize(n);
  c.des.resize(n);
  if (in.node_discs) {
    c.asc_disc.resize(n);
    c.des_disc.resize(n);
  }
  const int nc = (int)in.words.size();
  parallel_for(nc, [&](int k) {
    const uint32_t *p = in.words[k].data();
    for (uint32_t i = in.begin[k]; i < in.begin[k + 1]; ++i) {
      c.cell[i] = p[0];
      memcpy(&c.fn[i], &p[1], 4);
      c.pair[i] = p[2];
      c.flags[i] = (uint8_t)p[3];
      const uint32_t na = p[4], nd = p[5], nad = p[6], ndd = p[7];
      p += 8;
      for (uint32_t e = 0; e < na; ++e) c.asc[i].push_sorted(p[e]);
      p += na;
      for (uint32_t e = 0; e < nd; ++e) c.des[i].push_sorted(p[e]);
      p += nd;
      if (nad) c.asc_disc[i].assign(p, p + nad);
      p += nad;
      if (ndd) c.des_disc[i].assign(p, p + ndd);
      p += ndd;
    }
    std::vector<uint32_t>().swap(in.words[k]);
  });
  c.idx = std::move(in.idx);
  c.xd_off = std::move(in.xd_off);
  c.xd_cells = std::move(in.xd_cells);
  c.sd_off = std::move(in.sd_off);
  c.sd_cells = std::move(in.sd_cells);
  in.words.clear();
  in.begin.clear();
  in.n = 0;
  m = std::move(c);
}

// writeout_connectivity (grid_dataset.cpp:1858-1907) and
// addCriticalPointsToMSComplex (discreteMorseAlgorithm.h:244-268): nodes in
// raster order; each saddle connects to the owners of its facets (des) and
// cofacets (asc).  Arc multisets are order-free, so they are built in two
// passes (counts, then fills in index order, already sorted).
void complex_build(Complex &c, const Rect &rect, const Rect &ext, uint32_t K, const uint32_t *crit,
                   const float *fn, const uint32_t *conn4, const uint32_t *pair) {
  c.rect = rect;
  c.ext = ext;
  c.reserve_nodes(Complex::with_headroom(K));
  c.cell.assign(crit, crit + K);
  c.fn.assign(fn, fn + K);
  c.pair.assign(K, UINT32_MAX);
  c.flags.assign(K, 0);
  c.asc.assign(K, ConnSet());
  c.des.assign(K, ConnSet());
  c.idx.reserve(K);
  // the home slots are random: start loading them a few keys ahead
  constexpr uint32_t AHEAD = 16;
  for (uint32_t k = 0; k < K; ++k) {
    if (k + AHEAD < K) c.idx.prefetch(crit[k + AHEAD]);
    c.idx.insert(crit[k], k);
  }
  for (uint32_t k = 0; k < K; ++k) {
    if (pair[k] != UINT32_MAX) {
      c.flags[k] |= CPF_BCANC;
      c.pair[k] = pair[k];
    }
  }
  // saddle side: sorted small sets
  for (uint32_t k = 0; k < K; ++k) {
    if (cell_dim(crit[k]) != 1) continue;
    const uint32_t *q = &conn4[4 * (size_t)k];
    uint32_t d0 = q[0], d1 = q[1], a0 = q[2], a1 = q[3];
    if (d0 != UINT32_MAX && d1 != UINT32_MAX && d1 < d0) std::swap(d0, d1);
    if (d0 != UINT32_MAX) c.des[k].push_sorted(d0);
    if (d1 != UINT32_MAX) c.des[k].push_sorted(d1);
    if (a0 != UINT32_MAX && a1 != UINT32_MAX && a1 < a0) std::swap(a0, a1);
    if (a0 != UINT32_MAX) c.asc[k].push_sorted(a0);
    if (a1 != UINT32_MAX) c.asc[k].push_sorted(a1);
  }
  // extremum side: saddles visited in increasing index, so appends stay
  // sorted (the extrema's sets are random accesses: start loading those of
  // a saddle a few nodes ahead)
  for (uint32_t k = 0; k < K; ++k) {
    if (k + AHEAD < K && cell_dim(crit[k + AHEAD]) == 1) {
      for (uint32_t v : c.des[k + AHEAD]) __builtin_prefetch(&c.asc[v]);
      for (uint32_t v : c.asc[k + AHEAD]) __builtin_prefetch(&c.des[v]);
    }
    if (cell_dim(crit[k]) != 1) continue;
    for (uint32_t v : c.des[k]) {
      if (cell_dim(crit[v]) != 0) throw std::logic_error("must connect i,i+1 cp (or vice versa)");
      c.asc[v].push_sorted(k);
    }
    for (uint32_t v : c.asc[k]) {
      if (cell_dim(crit[v]) != 2) throw std::logic_error("must connect i,i+1 cp (or vice versa)");
      c.des[v].push_sorted(k);
    }
  }
}

void complex_from_tile(const mscx_tile *t, Complex &out) {
  std::vector<uint32_t> crit, conn, pr;
  std::vector<float> fn;
  tile_critical(t, crit, fn, conn, pr);
  Rect r{t->rect[0], t->rect[1], t->rect[2], t->rect[3]}, e{t->ext[0], t->ext[1], t->ext[2], t->ext[3]};
  complex_build(out, r, e, (uint32_t)crit.size(), crit.data(), fn.data(), conn.data(), pr.data());
}

static constexpr uint64_t kMaxArcProduct = 1ull << 20;
// Arc-multiset entries one simplify or un_simplify call may add (4 bytes
// each: about 1 GiB).  Legitimate heavy cases stay far below it: a noisy
// 1024^2 Gaussian sum simplified at 1% grows to 4.4e7 entries (165K
// distinct arcs with large multiplicities) after un_simplify.
static constexpr int64_t kArcEntryBudget = (int64_t)1 << 28;

// Heap pops one simplify call may make under a work limit, whatever the
// complex's size (a few seconds of work): large noisy grids, where each
// cancellation re-pushes tens of thousands of deferred pairs, reach it long
// before the per-entry bound.
static constexpr double kPopCap = (double)(1ull << 27);

// Arc-entry budget of one simplify / un_simplify call on m under a work
// limit (see simplify in complex.hpp).
static int64_t arc_entry_budget(const Complex &m, double work_limit) {
  if (!(work_limit > 0)) return kArcEntryBudget;
  size_t arcs = 0;
  for (uint32_t i = 0; i < m.size(); ++i) arcs += m.asc[i].size() + m.des[i].size();
  const double b = std::max((double)(1 << 24), 2.0 * work_limit * (double)arcs);
  return (int64_t)std::min(b, (double)kArcEntryBudget);
}

// cancelPairs (discreteMorseAlgorithm.h:348-464 / :466-585): cp0 is made
// the lower-dimensional cell; every asc(cp0) x des(cp1) pair is connected
// (duplicates kept, count >= 2 marks a strangulation), cp0/cp1 are removed
// from their neighbours, cp0.des and cp1.asc are cleared.
void cancel_pairs(Complex &m, uint32_t cp0, uint32_t cp1, std::vector<uint32_t> *new_edges) {
  if (!m.asc[cp0].contains(cp1)) std::swap(cp0, cp1);
  if (!m.asc[cp0].contains(cp1)) return;  // the reference logs and returns
  // asc[cp0] and des[cp1] are not modified inside the loops below (an
  // element equal to cp1 / cp0 is skipped), so iterating them directly is safe.
  const ConnSet &A = m.asc[cp0];
  const ConnSet &D = m.des[cp1];
  // Under dense cancellation (noisy or tie-heavy fields at large thresholds)
  // repeated strangulations multiply arc multiplicities at every step and
  // the reference grows without bound.  Fail the call instead of exhausting
  // host memory.
  const int64_t prod = (int64_t)A.size() * (int64_t)D.size();
  m.arc_budget -= prod;
  if (prod > (int64_t)kMaxArcProduct || m.arc_budget < 0)
    throw Error(MSCX_ENOMEM, "arc multiplicity explosion in cancelPairs (dense cancellation at this threshold)");
  for (uint32_t a : A) {
    for (uint32_t d : D) {
      if (a == cp1 || d == cp0) continue;
      m.des[a].insert(d);
      m.asc[d].insert(a);
      if (new_edges) {
        new_edges->push_back(d);
        new_edges->push_back(a);
      }
      if (m.des[a].count(d) >= 2) {
        m.flags[a] |= CPF_STRANGULATION;
        m.flags[d] |= CPF_STRANGULATION;
      }
    }
  }
  for (uint32_t a : A) {
    if (a == cp1) continue;
    m.des[a].erase_all(cp0);
  }
  for (uint32_t d : m.des[cp0]) m.asc[d].erase_all(cp0);
  for (uint32_t a : m.asc[cp1]) m.des[a].erase_all(cp1);
  for (uint32_t d : D) {
    if (d == cp0) continue;
    m.asc[d].erase_all(cp1);
  }
  m.flags[cp0] |= CPF_CANCELLED;
  m.des[cp0].clear();
  m.flags[cp1] |= CPF_CANCELLED;
  m.asc[cp1].clear();
}

// uncancel_pairs (discreteMorseAlgorithm.h:1006-1089).  The rebuilt
// multiset is gathered unsorted and sorted once (same contents as the
// reference's element-wise multiset inserts).
void uncancel_pairs(Complex &m, uint32_t cp1, uint32_t cp2) {
  if (!m.asc[cp1].contains(cp2)) std::swap(cp1, cp2);
  ConnSet *acdc[2] = {&m.asc[cp1], &m.des[cp2]};
  uint32_t idxs[2] = {cp1, cp2};
  std::vector<uint32_t> fresh;
  for (int i = 0; i < 2; ++i) {
    uint32_t other = idxs[(i + 1) % 2];
    if (!acdc[i]->contains(other)) throw std::logic_error("cancellable pair is not connected");
    fresh.clear();
    for (uint32_t v : *acdc[i]) {
      if (v == other) continue;
      if (!(m.flags[v] & CPF_BCANC)) {
        fresh.push_back(v);
        continue;
      }
      if (m.flags[v] & CPF_CANCELLED) throw std::logic_error("*acdc_conn_it should not have been cancelled yet");
      uint32_t pr = m.pair[v];
      if (m.pair[pr] != v) throw std::logic_error("*acdc_conn_it and its pair dont agree on pairing");
      const ConnSet &pc = (i == 0) ? m.asc[pr] : m.des[pr];
      for (uint32_t w : pc) {
        if (m.flags[w] & CPF_BCANC)
          throw std::logic_error("the connected cancelable cp shold not  contain a cancelable cp in its connections");
        fresh.push_back(w);
      }
    }
    m.arc_budget -= (int64_t)fresh.size();
    if (m.arc_budget < 0)
      throw Error(MSCX_ENOMEM, "arc multiplicity explosion in uncancel_pairs (dense cancellation at this threshold)");
    std::sort(fresh.begin(), fresh.end());
    acdc[i]->clear();
    for (uint32_t w : fresh) acdc[i]->push_sorted(w);
  }
  m.flags[cp1] &= (uint8_t)~CPF_CANCELLED;
  m.flags[cp2] &= (uint8_t)~CPF_CANCELLED;
}

void merge_ixn(const Complex &a, const Complex &b, Rect &ixn) { merge_ixn(a.rect, b.rect, ixn); }

void merge_ixn(const Rect &ra, const Rect &rb, Rect &ixn) {
  // rect_t::intersection (rectangle_complex.h:208-220), called as
  // msc2.m_rect.intersection(msc1.m_rect)
  const Rect &r1 = rb, &r2 = ra;
  bool inter = !(r1.x1 < r2.x0 || r2.x1 < r1.x0 || r1.y1 < r2.y0 || r2.y1 < r1.y0);
  if (!inter) throw std::logic_error("rects should intersect for merge");
  ixn = Rect{std::max(r2.x0, r1.x0), std::max(r2.y0, r1.y0), std::min(r2.x1, r1.x1), std::min(r2.y1, r1.y1)};
  if (ixn.x0 != ixn.x1 && ixn.y1 != ixn.y0) throw std::logic_error("rects must merge along a 1 manifold");
  if (ixn.x0 == ixn.x1 && ixn.y0 == ixn.y1) throw std::logic_error("rects cannot merge along a point alone");
}

// lexicographic (x, y) min / max of two points
static inline void pmin(int ax, int ay, int bx, int by, int &x, int &y) {
  if (ax < bx || (ax == bx && ay < by)) { x = ax; y = ay; } else { x = bx; y = by; }
}
static inline void pmax(int ax, int ay, int bx, int by, int &x, int &y) {
  if (ax < bx || (ax == bx && ay < by)) { x = bx; y = by; } else { x = ax; y = ay; }
}
static Rect rect_of_points(int ax, int ay, int bx, int by) {
  return Rect{std::min(ax, bx), std::min(ay, by), std::max(ax, bx), std::max(ay, by)};
}

// The union's rect and ext (grid_mscomplex.cpp:44-50): rect_t of the
// lexicographic min of the bottom-left corners and max of the top-right ones.
void merge_rects(const Rect &r1, const Rect &e1, const Rect &r2, const Rect &e2, Rect &r, Rect &e) {
  int bx, by, tx, ty;
  pmin(r1.x0, r1.y0, r2.x0, r2.y0, bx, by);
  pmax(r1.x1, r1.y1, r2.x1, r2.y1, tx, ty);
  r = rect_of_points(bx, by, tx, ty);
  pmin(e1.x0, e1.y0, e2.x0, e2.y0, bx, by);
  pmax(e1.x1, e1.y1, e2.x1, e2.y1, tx, ty);
  e = rect_of_points(bx, by, tx, ty);
}

// merge_up (grid_mscomplex.cpp:24-156)
void merge_up(const Complex &m1, const Complex &m2, Complex &res) {
  Rect ixn;
  merge_ixn(m1, m2, ixn);
  res = Complex();
  Complex *out = &res;
  merge_rects(m1.rect, m1.ext, m2.rect, m2.ext, out->rect, out->ext);
  const Complex *arr[2] = {&m1, &m2};

  // The reference walks msc1's then msc2's critical points, replicating
  // each one not yet in the union, then copies every arc.  Here the union's
  // indices are assigned first (at[i][j]: the index in `out` of m_i's node
  // j, UINT32_MAX for a cancelled node that is not replicated): msc1's
  // non-cancelled nodes in order, then msc2's nodes that are not msc1's
  // non-cancelled cells, in order -- the indices the reference's loop
  // gives.  The union's arrays are then filled by index; its cell map is
  // built on one thread while the others build the arc sets of disjoint
  // ranges of destination nodes (each set from its one or two source
  // nodes: a multiset's content does not depend on the order of its
  // inserts); the pair indices are looked up last, in the reference's order.
  const int nt = host_threads();
  const bool par = nt >= 2 && (size_t)m1.size() + m2.size() >= 16384;
  constexpr int CHUNK = 4096;
  auto chunks = [](uint32_t n) { return (int)((n + CHUNK - 1) / CHUNK); };
  std::vector<uint32_t> at[2];
  uint32_t n = 0;
  at[0].assign(m1.size(), UINT32_MAX);
  for (uint32_t j = 0; j < m1.size(); ++j)
    if (!(m1.flags[j] & CPF_CANCELLED)) at[0][j] = n++;
  const uint32_t n1 = n;
  // m2's cell among the union's cells so far, i.e. among m1's non-cancelled
  // ones (m1's map may also hold zero entries for cells it never held, hence
  // the check of the cell itself); the lookups are independent
  at[1].resize(m2.size());
  auto lookup = [&](int c) {
    const uint32_t b = (uint32_t)c * CHUNK, e = std::min<uint32_t>(m2.size(), b + CHUNK);
    for (uint32_t j = b; j < e; ++j) {
      if (j + 16 < e) m1.idx.prefetch(m2.cell[j + 16]);
      const uint32_t k = m1.idx.find(m2.cell[j]);
      at[1][j] = (k != UINT32_MAX && k < m1.size() && m1.cell[k] == m2.cell[j]) ? at[0][k] : UINT32_MAX;
    }
  };
  parallel_for(chunks(m2.size()), lookup, par ? nt : 1);
  std::vector<uint32_t> src[2];  // the source node of each destination, per complex
  src[0].resize(n1);
  for (uint32_t j = 0; j < m1.size(); ++j)
    if (at[0][j] != UINT32_MAX) src[0][at[0][j]] = j;
  std::vector<uint32_t> shared(n1, UINT32_MAX);  // m2's node on m1's cell d
  for (uint32_t j = 0; j < m2.size(); ++j) {
    if (at[1][j] != UINT32_MAX) {
      shared[at[1][j]] = j;
      continue;
    }
    if (m2.flags[j] & CPF_CANCELLED) continue;
    at[1][j] = n++;
  }
  src[1].resize(n - n1);
  for (uint32_t j = 0; j < m2.size(); ++j)
    if (at[1][j] >= n1 && at[1][j] != UINT32_MAX) src[1][at[1][j] - n1] = j;
  out->reserve_nodes(Complex::with_headroom(n));
  out->cell.resize(n);
  out->fn.resize(n);
  out->pair.assign(n, UINT32_MAX);
  out->flags.resize(n);
  out->asc.resize(n);
  out->des.resize(n);
  // m2's nodes on m1's cells keep m1's data
  auto fill = [&](int c) {
    const uint32_t b = (uint32_t)c * CHUNK, e = std::min<uint32_t>(n, b + CHUNK);
    for (uint32_t d = b; d < e; ++d) {
      const Complex &m = d < n1 ? m1 : m2;
      const uint32_t j = d < n1 ? src[0][d] : src[1][d - n1];
      out->cell[d] = m.cell[j];
      out->fn[d] = m.fn[j];
      out->flags[d] = m.flags[j];
    }
  };
  parallel_for(chunks(n), fill, par ? nt : 1);
  auto build_map = [&] {
    out->idx.reserve((size_t)m1.size() + m2.size());
    for (uint32_t d = 0; d < n; ++d) {
      if (d + 16 < n) out->idx.prefetch(out->cell[d + 16]);
      if (out->idx.find_or_insert(out->cell[d], d) != UINT32_MAX) throw std::logic_error("this cp is present in msc");
    }
  };
  // the non-cancelled arcs of m_i's node j into the sets of destination d;
  // m2's arcs between two cells of the shared line are dropped
  auto copy_node = [&](int i, uint32_t j, uint32_t d) {
    const Complex &m = *arr[i];
    if (m.flags[j] & CPF_CANCELLED) return;
    const uint32_t *ai = at[i].data();
    const bool cmn = (i == 1) && ixn.contains(cx_of(m.cell[j]), cy_of(m.cell[j]));
    for (int k = 0; k < 2; ++k) {
      const ConnSet &from = k == 0 ? m.asc[j] : m.des[j];
      ConnSet &dst = k == 0 ? out->asc[d] : out->des[d];
      for (uint32_t v : from) {
        if (cmn) {
          const uint32_t vc = m.cell[v];
          if (ixn.contains(cx_of(vc), cy_of(vc))) continue;
        }
        if (m.flags[v] & CPF_CANCELLED) continue;
        dst.insert(ai[v]);
      }
    }
  };
  auto arcs = [&](int c) {
    const uint32_t b = (uint32_t)c * CHUNK, e = std::min<uint32_t>(n, b + CHUNK);
    for (uint32_t d = b; d < e; ++d) {
      if (d < n1) {
        copy_node(0, src[0][d], d);
        if (shared[d] != UINT32_MAX) copy_node(1, shared[d], d);
      } else {
        copy_node(1, src[1][d - n1], d);
      }
    }
  };
  if (par) {
    std::exception_ptr em;
    std::thread tm([&] {
      try {
        build_map();
      } catch (...) {
        em = std::current_exception();
      }
    });
    try {
      parallel_for(chunks(n), arcs, nt - 1);
    } catch (...) {
      tm.join();
      throw;
    }
    tm.join();
    if (em) std::rethrow_exception(em);
  } else {
    build_map();
    for (int c = 0; c < chunks(n); ++c) arcs(c);
  }
  // map::operator[] on each boundary-cancellable node's pair cell (which
  // inserts a zero entry when the pair is absent, as the reference's lookup
  // does), m1's nodes then m2's, so that m2's lookup wins for a shared node.
  // The lookups of one complex run in parallel; the cells they miss get
  // their zero entries afterwards, in node order (a later miss of the same
  // cell then finds that entry: 0 either way).
  for (int i = 0; i < 2; ++i) {
    const Complex &m = *arr[i];
    const uint32_t *ai = at[i].data();
    const uint32_t nm = m.size();
    std::vector<std::vector<uint32_t>> miss(chunks(nm));
    auto look = [&](int c) {
      const uint32_t b = (uint32_t)c * CHUNK, e = std::min<uint32_t>(nm, b + CHUNK);
      for (uint32_t j = b; j < e; ++j) {
        if ((m.flags[j] & CPF_CANCELLED) || !(m.flags[j] & CPF_BCANC)) continue;
        const uint32_t v = out->idx.find(m.cell[m.pair[j]]);
        if (v != UINT32_MAX)
          out->pair[ai[j]] = v;
        else
          miss[c].push_back(j);
      }
    };
    parallel_for(chunks(nm), look, par ? nt : 1);
    for (const std::vector<uint32_t> &mc : miss)
      for (uint32_t j : mc) out->pair[ai[j]] = out->idx.get_or_zero(m.cell[m.pair[j]]);
  }
  merge_up_line(*out, ixn);
}

// The shared-line cancellations of merge_up (grid_mscomplex.cpp:124-153).
void merge_up_line(Complex &m, const Rect &ixn) {
  for (int y = ixn.y0; y <= ixn.y1; ++y) {
    for (int x = ixn.x0; x <= ixn.x1; ++x) {
      uint32_t c = (uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16);
      uint32_t si = m.idx.find(c);
      if (si == UINT32_MAX) throw std::logic_error("missing common bndry cp");
      if ((m.flags[si] & CPF_CANCELLED) || !(m.flags[si] & CPF_BCANC)) continue;
      uint32_t pi = m.pair[si];
      uint32_t p = m.cell[pi];
      if (!m.rect.interior(x, y) && !m.ext.on_boundary(x, y)) continue;
      if (!m.rect.interior(cx_of(p), cy_of(p)) && !m.ext.on_boundary(cx_of(p), cy_of(p))) continue;
      cancel_pairs(m, si, pi, nullptr);
    }
  }
}

// The shared-line uncancellations of merge_down (grid_mscomplex.cpp:172-202),
// in reverse raster order.
void merge_down_line(Complex &P, const Rect &ixn) {
  for (int y = ixn.y1; y >= ixn.y0; --y) {
    for (int x = ixn.x1; x >= ixn.x0; --x) {
      uint32_t c = (uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16);
      uint32_t si = P.idx.find(c);
      if (si == UINT32_MAX) throw std::logic_error("missing common bndry cp");
      if (!(P.flags[si] & CPF_CANCELLED)) continue;
      uint32_t pi = P.pair[si];
      uint32_t p = P.cell[pi];
      if (!P.rect.interior(x, y) && !P.ext.on_boundary(x, y)) continue;
      if (!P.rect.interior(cx_of(p), cy_of(p)) && !P.ext.on_boundary(cx_of(p), cy_of(p))) continue;
      uncancel_pairs(P, si, pi);
    }
  }
}

// merge_down (grid_mscomplex.cpp:158-322)
void merge_down(Complex &P, Complex &m1, Complex &m2) {
  Rect ixn;
  merge_ixn(m1, m2, ixn);
  merge_down_line(P, ixn);
  // The two children are updated from P (which is only read from here on)
  // independently of each other: large ones on two threads.  Within a child,
  // a first pass follows the reference's walk over the boundary-cancellable
  // nodes without copying arcs: it makes the replicas in the order the copies
  // would make them, sets the pair fields and raises the reference's checks.
  // The arc sets of every node present in the child are then rewritten from
  // the final index map, each by one thread (a set's content does not depend
  // on when its neighbours were replicated).  Should two of P's nodes map to
  // one node of the child (possible only through a zero entry of its cell
  // map), the sets are rewritten in the reference's order on one thread.
  auto down = [&](Complex &m, int nt) {
    std::vector<uint32_t> at;
    // at[j]: m's index of P's cell j (UINT32_MAX when absent), kept equal to
    // m.idx.find(P.cell[j]) as nodes are replicated into m
    at.resize(P.size());
    constexpr uint32_t AHEAD = 16, CHUNK = 8192;
    const uint32_t np = P.size();
    parallel_for(
        (int)((np + CHUNK - 1) / CHUNK),
        [&](int c) {
          const uint32_t b = (uint32_t)c * CHUNK, e = std::min(np, b + CHUNK);
          for (uint32_t j = b; j < e; ++j) {
            if (j + AHEAD < e) m.idx.prefetch(P.cell[j + AHEAD]);
            at[j] = m.idx.find(P.cell[j]);
          }
        },
        nt);
    auto rep = [&](uint32_t x) {
      const uint32_t r = m.replicate(P, x);
      m.fn.push_back(P.fn[x]);
      at[x] = r;
      return r;
    };
    std::vector<uint32_t> fill;  // P's nodes whose sets are rewritten in m, in the reference's order
    fill.reserve(P.size());
    for (uint32_t j = 0; j < P.size(); ++j) {
      if (P.flags[j] & CPF_CANCELLED) throw std::logic_error("all cps ought to be uncancelled by now");
      if (!(P.flags[j] & CPF_BCANC)) continue;
      uint32_t sp = P.pair[j];
      bool src_in = at[j] != UINT32_MAX;
      bool sp_in = at[sp] != UINT32_MAX;
      if (!src_in && !sp_in) continue;
      if (!src_in) rep(j);
      if (!sp_in) rep(sp);
      uint32_t d = at[j];
      if (!src_in || !sp_in || !(m.flags[d] & CPF_BCANC)) {
        uint32_t dp = at[sp];
        m.flags[d] |= CPF_BCANC;
        m.flags[dp] |= CPF_BCANC;
        m.pair[d] = dp;
        m.pair[dp] = d;
      }
      fill.push_back(j);
      for (int k = 0; k < 2; ++k)
        for (uint32_t v : (k == 0 ? P.asc[j] : P.des[j])) {
          if (P.flags[v] & CPF_BCANC) throw std::logic_error("only non cancellable cps must be remaining");
          if (at[v] == UINT32_MAX) rep(v);
        }
    }
    for (uint32_t j = 0; j < P.size(); ++j)
      if (!(P.flags[j] & CPF_BCANC) && at[j] != UINT32_MAX) fill.push_back(j);
    // a boundary-cancellable node's neighbours are all in m (replicated
    // above); another node's neighbours absent from m are dropped
    auto rewrite = [&](uint32_t j) {
      const uint32_t d = at[j];
      for (int k = 0; k < 2; ++k) {
        const ConnSet &src = k == 0 ? P.asc[j] : P.des[j];
        ConnSet &dst = k == 0 ? m.asc[d] : m.des[d];
        dst.clear();
        for (uint32_t v : src) {
          const uint32_t mv = at[v];
          if (mv != UINT32_MAX) dst.insert(mv);
        }
      }
    };
    bool distinct = true;
    if (nt > 1 && fill.size() >= 8192) {
      std::vector<uint8_t> seen(m.size(), 0);
      for (uint32_t j : fill) {
        if (seen[at[j]]) {
          distinct = false;
          break;
        }
        seen[at[j]] = 1;
      }
    }
    if (nt > 1 && fill.size() >= 8192 && distinct) {
      // rewrite() throws only when an allocation fails
      const size_t nf = fill.size(), chunk = 512;
      parallel_for(
          (int)((nf + chunk - 1) / chunk),
          [&](int c) {
            for (size_t q = (size_t)c * chunk; q < std::min(nf, (size_t)(c + 1) * chunk); ++q) rewrite(fill[q]);
          },
          nt);
    } else {
      for (uint32_t j : fill) rewrite(j);
    }
  };
  const int nt = host_threads();
  if (P.size() >= 4096 && nt > 1) {
    const int per = std::max(1, nt / 2);
    std::exception_ptr err;
    std::thread other([&] {
      try {
        down(m2, per);
      } catch (...) {
        err = std::current_exception();
      }
    });
    try {
      down(m1, per);
    } catch (...) {
      other.join();
      throw;
    }
    other.join();
    if (err) std::rethrow_exception(err);
  } else {
    down(m1, 1);
    down(m2, 1);
  }
}

// persistence_comparator_t (grid_mscomplex.cpp:331-379).  Its geometric
// tie-break compares the first pair with itself, so two pairs of equal
// persistence are equivalent; only |f2 - f1| orders the heap.  The heap
// element carries that persistence (the function values do not change while
// simplifying), so a comparison reads no other memory; the sequence of heap
// operations, and hence the order among equal persistences, is the one
// std::priority_queue produces for the reference's (uint, uint) elements.
// Heap element: the persistence as the bit pattern of a non-negative float
// (ordered like the float) and the slot of the (first, second) pair in a
// side array.  ExactHeap reproduces std::priority_queue's element moves, so
// equal persistences leave in the reference's order.

// simplify (grid_mscomplex.cpp:384-512)
void simplify(Complex &m, std::vector<std::pair<uint32_t, uint32_t>> &canc, double thresh, double work_limit) {
  typedef std::pair<uint32_t, uint32_t> P2;
  const float *fn = m.fn.data();
  std::vector<P2> slots;  // (first, second) of every pushed element
  auto mk = [&](uint32_t a, uint32_t b) {
    float p = std::abs(fn[b] - fn[a]);
    uint32_t key;
    memcpy(&key, &p, 4);
    slots.emplace_back(a, b);
    return ExactHeap::make(key, (uint32_t)(slots.size() - 1));
  };
  size_t narcs = 0;
  for (uint32_t i = 0; i < m.size(); ++i) narcs += m.des[i].size();
  m.arc_budget = arc_entry_budget(m, work_limit);
  m.pop_budget = work_limit > 0 ? (int64_t)std::min(kPopCap, std::max((double)(1 << 24), work_limit * (double)narcs))
                                 : INT64_MAX;
  ExactHeap pq;
  pq.reserve(narcs * 2 + 16);
  // the pops' random accesses span these two arrays: huge pages keep them
  // within reach of the TLB
  slots.reserve(narcs * 4 + 16);
  advise_huge(slots.data(), slots.capacity() * sizeof(P2));

  float max_val = std::numeric_limits<float>::min();
  float min_val = std::numeric_limits<float>::max();
  for (uint32_t i = 0; i < m.size(); ++i) {
    if (cell_dim(m.cell[i]) == 1) {
      for (int j = 0; j < 2; ++j) {
        const ConnSet &s = j == 0 ? m.asc[i] : m.des[i];
        if (s.size() != 2) continue;
        if (s[0] == s[1]) m.flags[i] |= CPF_STRANGULATION;
      }
    }
    max_val = std::max(max_val, m.fn[i]);
    min_val = std::min(min_val, m.fn[i]);
    for (uint32_t d : m.des[i]) pq.push(mk(i, d));
  }
  std::vector<uint64_t> resubmit;
  float max_persistence = max_val - min_val;
  std::vector<uint32_t> new_edges;
  const Rect &R = m.rect;
  // per node: dimension (bits 0-1) and rect-boundary flag (bit 2), so that
  // a popped pair is classified from two bytes instead of its cell ids
  std::vector<uint8_t> info(m.size());
  for (uint32_t i = 0; i < m.size(); ++i) {
    const uint32_t c = m.cell[i];
    info[i] = (uint8_t)(cell_dim(c) | (R.on_boundary(cx_of(c), cy_of(c)) ? 4u : 0u));
  }
  const uint8_t *flags = m.flags.data();

  while (!pq.empty()) {
    uint64_t hp = pq.top();
    const P2 cp = slots[ExactHeap::payload(hp)];
    pq.pop();
    if (--m.pop_budget < 0)
      throw Error(MSCX_ELIMIT, "simplification work limit reached (strangulation resubmissions at this threshold)");
    // the next element's pair is a random access into a large array
    if (!pq.empty()) __builtin_prefetch(&slots[ExactHeap::payload(pq.top())]);
    // the key is the persistence |fn[v2] - fn[v1]| computed at the push
    const uint32_t kb = ExactHeap::key(hp);
    float pers;
    memcpy(&pers, &kb, 4);
    if ((double)pers / (double)max_persistence > thresh) break;
    uint32_t v1 = cp.first, v2 = cp.second;
    if ((flags[v1] | flags[v2]) & CPF_CANCELLED) continue;
    uint8_t i1 = info[v1], i2 = info[v2];
    if ((i2 & 3u) == 1) {
      std::swap(v1, v2);
      std::swap(i1, i2);
    }
    const uint32_t d2 = i2 & 3u;
    const bool b1 = i1 & 4u, b2 = i2 & 4u;
    if (b1 != b2) continue;
    const ConnSet &c2 = (d2 / 2 == 0) ? m.asc[v2] : m.des[v2];
    bool strang = flags[v1] & CPF_STRANGULATION;
    if (strang && c2.size() != 1) {
      resubmit.push_back(hp);
      continue;
    }
    if (strang && b1) continue;
    new_edges.clear();
    cancel_pairs(m, v1, v2, &new_edges);
    m.flags[v1] |= CPF_BCANC;
    m.flags[v2] |= CPF_BCANC;
    m.pair[v1] = v2;
    m.pair[v2] = v1;
    canc.push_back(cp);
    for (size_t i = 0; i < new_edges.size(); i += 2) pq.push(mk(new_edges[i], new_edges[i + 1]));
    for (uint64_t r : resubmit) pq.push(r);
    resubmit.clear();
  }
}

void un_simplify(Complex &m, const std::vector<std::pair<uint32_t, uint32_t>> &canc, double work_limit) {
  m.arc_budget = arc_entry_budget(m, work_limit);
  for (auto it = canc.rbegin(); it != canc.rend(); ++it) uncancel_pairs(m, it->first, it->second);
  m.arc_budget = INT64_MAX;
}

// ---------------------------------------------------------------- discs
// postMergeFillDiscs (grid_dataset.cpp:1044-1231) in four parts:
//   disc_prepare  -- the pass over critical points (:1052-1140): owner
//                    redirects of boundary-cancellable saddles, seeds of
//                    saddle discs, the surviving-extremum maps
//   labels        -- per interior vertex / quad, the surviving extremum
//                    that owns it after one redirect (:1142-1187); computed
//                    on the GPU by k_disc_labels, or on the host here
//   disc_group    -- extremum discs as raster-ordered lists (CSR)
//   disc_paths    -- saddle 1-manifolds traced from the seeds (:1189-1228)

static inline uint32_t pack_xy(int x, int y) { return (uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16); }

uint32_t surviving_extremum(const Complex &m, uint32_t c, uint32_t dim) {
  const uint32_t i = m.idx.find(c);
  // (a host-merged complex's map may hold zero entries for cells it lacks)
  if (i == UINT32_MAX || i >= m.size() || m.cell[i] != c) return UINT32_MAX;
  return (cell_dim(c) == dim && !(m.flags[i] & CPF_BCANC)) ? i : UINT32_MAX;
}

void disc_prepare(Complex &m, const Rect &rect, const CodesView &cv, DiscPrep &out, bool with_maps) {
  auto in_r = [&](uint32_t c) { return c != UINT32_MAX && rect.contains(cx_of(c), cy_of(c)); };
  CellMap &redir = out.redir;  // cell -> slot in out.redirect (later assignments overwrite)
  redir = CellMap();
  if (with_maps) redir.reserve(m.size() / 4 + 16);
  out.redirect.clear();
  out.redir_node.assign(m.size(), UINT32_MAX);
  out.minmap.clear();
  out.maxmap.clear();
  out.surv_saddles.clear();
  // node `key` (cell m.cell[key]) redirected to node `to`
  auto set_redirect = [&](uint32_t key, uint32_t to) {
    out.redir_node[key] = to;
    if (!with_maps) return;
    const uint32_t cell = m.cell[key], val = m.cell[to];
    uint32_t s = redir.find(cell);
    if (s == UINT32_MAX) {
      redir.set(cell, (uint32_t)out.redirect.size());
      out.redirect.emplace_back(cell, val);
    } else {
      out.redirect[s].second = val;
    }
  };
  CellMap seen_min, seen_max;
  // the discs' initial entries, in the order the reference appends them
  // (node order; a saddle's own seeds when its node comes): counted in the
  // first pass, written in the second
  const size_t nslot = 2 * (size_t)m.size();
  out.init_off.assign(nslot + 1, 0);
  out.traced.assign(nslot, 0);
  // the nodes that append entries (boundary-cancellable extrema, surviving
  // saddles), in node order: the second pass visits only them
  std::vector<uint32_t> pushers;
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<uint32_t> cur;
    if (pass == 1) {
      for (size_t k = 0; k < nslot; ++k) out.init_off[k + 1] += out.init_off[k];
      out.init_cells.resize(out.init_off[nslot]);
      cur.assign(out.init_off.begin(), out.init_off.end() - 1);
    }
    auto push = [&](uint32_t node, int j, uint32_t cell) {
      const size_t sl = 2 * (size_t)node + j;
      if (pass == 0)
        out.init_off[sl + 1]++;
      else
        out.init_cells[cur[sl]++] = cell;
    };
    const uint32_t nvisit = pass == 0 ? m.size() : (uint32_t)pushers.size();
    for (uint32_t v = 0; v < nvisit; ++v) {
      const uint32_t i = pass == 0 ? v : pushers[v];
      uint32_t c = m.cell[i];
      uint32_t dim = cell_dim(c);
      if (pass == 0 && ((m.flags[i] & CPF_BCANC) != 0) == (dim != 1)) pushers.push_back(i);
      if (m.flags[i] & CPF_BCANC) {
        uint32_t pi = m.pair[i];
        uint32_t pc = m.cell[pi];
        if (dim == 1) {
          if (pass == 1) continue;
          const ConnSet &conn = (cell_dim(pc) == 0) ? m.des[i] : m.asc[i];
          if (conn.size() != 1) throw std::logic_error("I should be connected to exactly one surv extrema");
          if (m.pair[pi] != i) throw std::logic_error("My pair does not know Im paired with him");
          if (in_r(c)) set_redirect(i, conn[0]);
          if (in_r(pc)) set_redirect(pi, conn[0]);
        } else if (dim == 2) {
          for (uint32_t s : m.des[i])
            if (in_r(pc)) push(s, 0, pc);
        } else {
          for (uint32_t s : m.asc[i])
            if (in_r(pc)) push(s, 1, pc);
        }
      } else {
        if (dim == 0) {
          if (pass == 0 && with_maps && !seen_min.contains(c)) {  // map::insert keeps the first
            seen_min.set(c, i);
            out.minmap.emplace_back(c, i);
          }
        } else if (dim == 1) {
          int x = cx_of(c), y = cy_of(c);
          const int inc[4][2] = {{x, y + 1}, {x, y - 1}, {x - 1, y}, {x + 1, y}};
          for (int j = 0; j < 4; ++j) {
            int ix = inc[j][0], iy = inc[j][1];
            if (!rect.contains(ix, iy)) continue;
            uint8_t code = cv.at(ix, iy);
            if ((code & CF_PAIRED) && !(code & CF_CRITICAL)) {
              int d = code_dir(code);
              uint32_t pr = pack_xy(ix + dir_dx(d), iy + dir_dy(d));
              push(i, ((ix & 1) + (iy & 1)) == 0 ? 1 : 0, pr);
            }
          }
          if (pass == 0) {
            out.surv_saddles.push_back(i);
            out.traced[2 * (size_t)i] = out.traced[2 * (size_t)i + 1] = 1;
          }
        } else {
          if (pass == 0 && with_maps && !seen_max.contains(c)) {
            seen_max.set(c, i);
            out.maxmap.emplace_back(c, i);
          }
        }
      }
    }
  }
  if (!with_maps) return;
  // sorted: the slots in redir move with their entries
  std::sort(out.redirect.begin(), out.redirect.end());
  for (uint32_t k = 0; k < out.redirect.size(); ++k) redir.set(out.redirect[k].first, k);
  std::sort(out.minmap.begin(), out.minmap.end());
  std::sort(out.maxmap.begin(), out.maxmap.end());
}

static inline uint32_t sorted_find(const std::vector<std::pair<uint32_t, uint32_t>> &v, uint32_t key) {
  auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(key, (uint32_t)0));
  return (it != v.end() && it->first == key) ? it->second : UINT32_MAX;
}

void disc_labels_host(const DiscPrep &dp, const Rect &rect, const Rect &ext, const uint32_t *own, int64_t own_pitch,
                      std::vector<uint32_t> &lv, std::vector<uint32_t> &lq) {
  auto in_r = [&](uint32_t c) { return c != UINT32_MAX && rect.contains(cx_of(c), cy_of(c)); };
  auto own2 = [&](uint32_t c) {
    uint32_t r = sorted_find(dp.redirect, c);
    return r != UINT32_MAX ? r : own[(int64_t)(cy_of(c) - ext.y0) * own_pitch + (cx_of(c) - ext.x0)];
  };
  auto label = [&](uint32_t c, const std::vector<std::pair<uint32_t, uint32_t>> &mp) -> uint32_t {
    uint32_t o = own2(c);
    if (in_r(o)) o = own2(o);
    if (cx_of(o) == -1 || cy_of(o) == -1) return UINT32_MAX;
    uint32_t k = sorted_find(mp, o);
    if (k == UINT32_MAX) throw std::logic_error("owner extremum not in surviving map");
    return k;
  };
  lv.clear();
  lq.clear();
  for (int y = rect.y0; y <= rect.y1; y += 2)
    for (int x = rect.x0; x <= rect.x1; x += 2) lv.push_back(label(pack_xy(x, y), dp.minmap));
  for (int y = rect.y0 + 1; y < rect.y1; y += 2)
    for (int x = rect.x0 + 1; x < rect.x1; x += 2) lq.push_back(label(pack_xy(x, y), dp.maxmap));
}

void disc_group(Complex &m, const Rect &rect, const uint32_t *lv, const uint32_t *lq) {
  const int nvx = (rect.x1 - rect.x0) / 2 + 1, nvy = (rect.y1 - rect.y0) / 2 + 1;
  const size_t nv = (size_t)nvx * nvy, nq = (size_t)(nvx - 1) * (nvy - 1);
  const uint32_t n = m.size();
  std::vector<uint32_t> &off = m.xd_off;
  off.assign((size_t)n + 1, 0);
  for (size_t r = 0; r < nv; ++r)
    if (lv[r] != UINT32_MAX) off[lv[r] + 1]++;
  for (size_t r = 0; r < nq; ++r)
    if (lq[r] != UINT32_MAX) off[lq[r] + 1]++;
  for (uint32_t i = 0; i < n; ++i) off[i + 1] += off[i];
  m.xd_cells.resize(off[n]);
  std::vector<uint32_t> pos(off.begin(), off.end() - 1);
  size_t r = 0;
  for (int y = rect.y0; y <= rect.y1; y += 2)
    for (int x = rect.x0; x <= rect.x1; x += 2, ++r)
      if (lv[r] != UINT32_MAX) m.xd_cells[pos[lv[r]]++] = pack_xy(x, y);
  r = 0;
  for (int y = rect.y0 + 1; y < rect.y1; y += 2)
    for (int x = rect.x0 + 1; x < rect.x1; x += 2, ++r)
      if (lq[r] != UINT32_MAX) m.xd_cells[pos[lq[r]]++] = pack_xy(x, y);
}

// The successors of one path cell (grid_dataset.cpp:1200-1224): its
// cofacets inside the ext (j = 0) or facets (j = 1), in the reference's
// order, that are not critical and lie in the rect, each replaced by its
// pair unless that is the path cell itself.
static inline int disc_successors(uint32_t pcell, int j, const Rect &rect, const Rect &ext, const CodesView &cv,
                                  uint32_t succ[4]) {
  const int px = cx_of(pcell), py = cy_of(pcell);
  uint32_t cets[4];
  int ct = 0;
  const int dx = px & 1, dy = py & 1;
  if (!ext.contains(px, py)) return 0;  // no path cell lies outside the ext
  if (j == 0) {  // cofacets, filtered to ext
    if (dx + dy == 0) {
      const int c4[4][2] = {{px, py + 1}, {px, py - 1}, {px - 1, py}, {px + 1, py}};
      for (int q = 0; q < 4; ++q)
        if (ext.contains(c4[q][0], c4[q][1])) cets[ct++] = pack_xy(c4[q][0], c4[q][1]);
    } else if (dx + dy == 1) {
      if (ext.contains(px + dy, py + dx)) cets[ct++] = pack_xy(px + dy, py + dx);
      if (ext.contains(px - dy, py - dx)) cets[ct++] = pack_xy(px - dy, py - dx);
    }
  } else {  // facets
    if (dx + dy == 1) {
      cets[ct++] = pack_xy(px + dx, py + dy);
      cets[ct++] = pack_xy(px - dx, py - dy);
    } else if (dx + dy == 2) {
      const int f4[4][2] = {{px, py + 1}, {px, py - 1}, {px - 1, py}, {px + 1, py}};
      for (int q = 0; q < 4; ++q) cets[ct++] = pack_xy(f4[q][0], f4[q][1]);
    }
  }
  int ns = 0;
  for (int q = 0; q < ct; ++q) {
    const int qx = cx_of(cets[q]), qy = cy_of(cets[q]);
    const uint8_t code = cv.at(qx, qy);
    if (code & CF_CRITICAL) continue;
    if (!rect.contains(qx, qy)) continue;
    const int d = code_dir(code);
    const uint32_t p = d ? pack_xy(qx + dir_dx(d), qy + dir_dy(d)) : UINT32_MAX;
    if (p == pcell) continue;
    succ[ns++] = p;
  }
  return ns;
}

void disc_trace_one(const uint32_t *init, size_t n_init, int j, const Rect &rect, const Rect &ext,
                    const CodesView &cv, std::vector<uint32_t> &out) {
  const size_t base = out.size();
  out.insert(out.end(), init, init + n_init);
  for (size_t k = base; k != out.size(); ++k) {
    uint32_t succ[4];
    const int ns = disc_successors(out[k], j, rect, ext, cv, succ);
    for (int q = 0; q < ns; ++q) out.push_back(succ[q]);
  }
}

void disc_trace_host(Complex &m, const DiscPrep &dp, const Rect &rect, const Rect &ext, const CodesView &cv) {
  const size_t nslot = 2 * (size_t)m.size();
  m.sd_off.assign(nslot + 1, 0);
  m.sd_cells.clear();
  for (size_t sl = 0; sl < nslot; ++sl) {
    const uint32_t *init = dp.init_cells.data() + dp.init_off[sl];
    const size_t n = dp.init_off[sl + 1] - dp.init_off[sl];
    if (dp.traced[sl] && n)
      disc_trace_one(init, n, (int)(sl & 1), rect, ext, cv, m.sd_cells);
    else
      m.sd_cells.insert(m.sd_cells.end(), init, init + n);
    m.sd_off[sl + 1] = (uint32_t)m.sd_cells.size();
  }
}

void fill_discs(Complex &m, const Rect &rect, const Rect &ext, const uint8_t *codes, std::vector<uint32_t> &own) {
  const int cw = ext.x1 - ext.x0 + 1;
  CodesView cv{codes, cw, ext.x0, ext.y0};
  DiscPrep dp;
  disc_prepare(m, rect, cv, dp);
  std::vector<uint32_t> lv, lq;
  disc_labels_host(dp, rect, ext, own.data(), cw, lv, lq);
  disc_group(m, rect, lv.data(), lq.data());
  disc_trace_host(m, dp, rect, ext, cv);
  for (const auto &r : dp.redirect) own[(size_t)(cy_of(r.first) - ext.y0) * cw + (cx_of(r.first) - ext.x0)] = r.second;
}

// write_discs (grid_mscomplex.cpp:542-590): raw packed ids appended to
// <prefix>asc_(x,y) and <prefix>des_(x,y).
void write_discs(const Complex &m, const std::string &prefix) {
  for (uint32_t i = 0; i < m.size(); ++i) {
    if (m.flags[i] & CPF_BCANC) continue;
    for (int j = 0; j < 2; ++j) {
      DiscSpan d = m.disc(i, j);
      if (d.n == 0) continue;
      std::ostringstream ss;
      ss << prefix << (j == 0 ? "asc_" : "des_") << "(" << cx_of(m.cell[i]) << "," << cy_of(m.cell[i]) << ")";
      std::ofstream os(ss.str(), std::ios::out | std::ios::ate | std::ios::app | std::ios::binary);
      if (!os.is_open()) throw Error(MSCX_EINVAL, "asc/des disc file not writeable");
      os.write((const char *)d.p, (std::streamsize)(d.n * 4));
    }
  }
}

// A host archive of one complex, the counterpart of the reference's
// Boost binary archives of a GridMSComplex (write_msgraph_to_archive /
// read_msgraph_from_archive, grid_datamanager.cpp:91-119; the fields of
// grid_mscomplex.cpp:604-644).  Boost archives are not portable, so the
// format is this library's own: an 8-byte magic, a version word, a reserved
// word, the word count, the complex's dump words (rect, ext, and per
// critical point its cell, value bits, pair, flags, arcs and discs) and an
// FNV-1a checksum of those words, all little-endian.  Discs are kept too,
// so a loaded complex dumps word for word as the saved one.
static const char kArchiveMagic[8] = {'M', 'S', 'C', 'X', 'C', 'P', 'X', '1'};
static const uint32_t kArchiveVersion = 1;

static uint64_t fnv1a(const uint32_t *w, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i)
    for (int b = 0; b < 4; ++b) {
      h ^= (w[i] >> (8 * b)) & 0xFFu;
      h *= 1099511628211ull;
    }
  return h;
}

void save_complex(const Complex &m, const std::string &path) {
  std::vector<uint32_t> w;
  m.dump(w);
  const uint32_t hdr[2] = {kArchiveVersion, 0};
  const uint64_t n = w.size(), sum = fnv1a(w.data(), w.size());
  std::ofstream os(path, std::ios::out | std::ios::trunc | std::ios::binary);
  if (!os.is_open()) throw Error(MSCX_EIO, "cannot write complex archive: " + path);
  os.write(kArchiveMagic, 8);
  os.write((const char *)hdr, sizeof hdr);
  os.write((const char *)&n, 8);
  os.write((const char *)w.data(), (std::streamsize)(w.size() * 4));
  os.write((const char *)&sum, 8);
  os.flush();
  if (!os.good()) throw Error(MSCX_EIO, "cannot write complex archive: " + path);
}

// A complex from its dump words (Complex::dump); every length is checked
// against the words left.  `what` names the source in the error.
void complex_from_words(const uint32_t *w, size_t n, Complex &m, const std::string &what) {
  auto bad = [&]() { return Error(MSCX_EIO, "corrupt complex archive: " + what); };
  size_t at = 0;
  auto take = [&](size_t k) {
    if (n - at < k) throw bad();
    const uint32_t *p = w + at;
    at += k;
    return p;
  };
  const uint32_t *h = take(9);
  const uint32_t ncp = h[0];
  Complex c;
  c.rect = Rect{(int)h[1], (int)h[2], (int)h[3], (int)h[4]};
  c.ext = Rect{(int)h[5], (int)h[6], (int)h[7], (int)h[8]};
  if ((uint64_t)ncp * 8 > n) throw bad();
  c.reserve_nodes(Complex::with_headroom(ncp));
  c.cell.resize(ncp);
  c.fn.resize(ncp);
  c.pair.resize(ncp);
  c.flags.resize(ncp);
  c.asc.resize(ncp);
  c.des.resize(ncp);
  c.idx.reserve(ncp);
  for (uint32_t i = 0; i < ncp; ++i) {
    const uint32_t *r = take(8);
    c.cell[i] = r[0];
    memcpy(&c.fn[i], &r[1], 4);
    c.pair[i] = r[2];
    if (r[3] > 0xFFu || (r[2] != UINT32_MAX && r[2] >= ncp)) throw bad();
    c.flags[i] = (uint8_t)r[3];
    const uint32_t na = r[4], nd = r[5], nad = r[6], ndd = r[7];
    for (int k = 0; k < 2; ++k) {
      const uint32_t cnt = k == 0 ? na : nd;
      const uint32_t *v = take(cnt);
      ConnSet &dst = k == 0 ? c.asc[i] : c.des[i];
      for (uint32_t e = 0; e < cnt; ++e) {
        if (v[e] >= ncp || (e && v[e] < v[e - 1])) throw bad();
        dst.push_sorted(v[e]);
      }
    }
    const uint32_t *ad = take(nad);
    const uint32_t *dd = take(ndd);
    if ((nad || ndd) && c.asc_disc.empty()) {
      c.asc_disc.resize(ncp);
      c.des_disc.resize(ncp);
    }
    if (nad) c.asc_disc[i].assign(ad, ad + nad);
    if (ndd) c.des_disc[i].assign(dd, dd + ndd);
    if (c.idx.find_or_insert(c.cell[i], i) != UINT32_MAX) throw bad();
  }
  if (at != n) throw bad();
  m = std::move(c);
}

void load_complex(const std::string &path, Complex &m) {
  std::ifstream is(path, std::ios::in | std::ios::binary);
  if (!is.is_open()) throw Error(MSCX_EIO, "cannot read complex archive: " + path);
  char magic[8];
  uint32_t hdr[2];
  uint64_t n = 0;
  is.read(magic, 8);
  is.read((char *)hdr, sizeof hdr);
  is.read((char *)&n, 8);
  if (!is.good() || memcmp(magic, kArchiveMagic, 8) != 0) throw Error(MSCX_EIO, "not a complex archive: " + path);
  if (hdr[0] != kArchiveVersion) throw Error(MSCX_EIO, "unsupported complex archive version: " + path);
  is.seekg(0, std::ios::end);
  const uint64_t bytes = (uint64_t)is.tellg();
  if (n < 9 || bytes != 24 + 4 * n + 8) throw Error(MSCX_EIO, "truncated complex archive: " + path);
  is.seekg(24);
  std::vector<uint32_t> w((size_t)n);
  uint64_t sum = 0;
  is.read((char *)w.data(), (std::streamsize)(4 * n));
  is.read((char *)&sum, 8);
  if (!is.good() || sum != fnv1a(w.data(), w.size())) throw Error(MSCX_EIO, "corrupt complex archive: " + path);

  complex_from_words(w.data(), w.size(), m, path);
}

}  // namespace mscx

using namespace mscx;

extern "C" {

int mscx_complex_from_tile(const mscx_tile *t, mscx_complex **out) {
  return guarded([&] {
    MSCX_REQUIRE(t && out, MSCX_EINVAL, "null argument");
    std::unique_ptr<mscx_complex> h(new mscx_complex);
    complex_from_tile(t, h->c);
    *out = h.release();
  });
}

int mscx_complex_from_critical(const int32_t rect[4], const int32_t ext[4], uint32_t K, const uint32_t *crit,
                               const float *fn, const uint32_t *conn4, const uint32_t *pair_idx,
                               mscx_complex **out) {
  return guarded([&] {
    MSCX_REQUIRE(rect && ext && out && (K == 0 || (crit && fn && conn4 && pair_idx)), MSCX_EINVAL,
                 "null argument");
    for (uint32_t k = 0; k < K; ++k) {
      for (int q = 0; q < 4; ++q)
        MSCX_REQUIRE(conn4[4 * (size_t)k + q] == UINT32_MAX || conn4[4 * (size_t)k + q] < K, MSCX_EINVAL,
                     "incidence index out of range");
      MSCX_REQUIRE(pair_idx[k] == UINT32_MAX || pair_idx[k] < K, MSCX_EINVAL, "pair index out of range");
    }
    std::unique_ptr<mscx_complex> h(new mscx_complex);
    Rect r{rect[0], rect[1], rect[2], rect[3]}, e{ext[0], ext[1], ext[2], ext[3]};
    complex_build(h->c, r, e, K, crit, fn, conn4, pair_idx);
    *out = h.release();
  });
}

int mscx_complex_fill_discs_host(mscx_complex *c, const int32_t rect[4], const int32_t ext[4], const uint8_t *codes,
                                 uint32_t *own) {
  return guarded([&] {
    MSCX_REQUIRE(c && rect && ext && codes && own, MSCX_EINVAL, "null argument");
    Rect r{rect[0], rect[1], rect[2], rect[3]}, e{ext[0], ext[1], ext[2], ext[3]};
    size_t n = (size_t)(e.x1 - e.x0 + 1) * (e.y1 - e.y0 + 1);
    std::vector<uint32_t> o(own, own + n);
    fill_discs(c->c, r, e, codes, o);
    memcpy(own, o.data(), n * 4);
  });
}

int mscx_complex_merge_up(const mscx_complex *a, const mscx_complex *b, mscx_complex **out) {
  return guarded([&] {
    MSCX_REQUIRE(a && b && out, MSCX_EINVAL, "null argument");
    std::unique_ptr<mscx_complex> h(new mscx_complex);
    merge_up(a->c, b->c, h->c);
    *out = h.release();
  });
}

int mscx_complex_merge_down(mscx_complex *p, mscx_complex *a, mscx_complex *b) {
  return guarded([&] {
    MSCX_REQUIRE(p && a && b, MSCX_EINVAL, "null argument");
    merge_down(p->c, a->c, b->c);
  });
}

int mscx_complex_simplify_un_simplify(mscx_complex *c, double threshold) {
  return guarded([&] {
    MSCX_REQUIRE(c, MSCX_EINVAL, "null argument");
    std::vector<std::pair<uint32_t, uint32_t>> canc;
    simplify(c->c, canc, threshold);
    un_simplify(c->c, canc);
  });
}

int mscx_tile_fill_discs(mscx_tile *t, mscx_complex *c) {
  return guarded([&] {
    MSCX_REQUIRE(t && c, MSCX_EINVAL, "null argument");
    Complex *cc = &c->c;
    fill_discs_device(&t, &cc, 1);
  });
}

int mscx_complex_write_discs(const mscx_complex *c, const char *prefix) {
  return guarded([&] {
    MSCX_REQUIRE(c && prefix, MSCX_EINVAL, "null argument");
    write_discs(c->c, prefix);
  });
}

int mscx_complex_save(const mscx_complex *c, const char *path) {
  return guarded([&] {
    MSCX_REQUIRE(c && path, MSCX_EINVAL, "null argument");
    save_complex(c->c, path);
  });
}

int mscx_complex_load(const char *path, mscx_complex **out) {
  return guarded([&] {
    MSCX_REQUIRE(path && out, MSCX_EINVAL, "null argument");
    std::unique_ptr<mscx_complex> h(new mscx_complex);
    load_complex(path, h->c);
    *out = h.release();
  });
}

int mscx_complex_dump(const mscx_complex *c, uint32_t *buf, size_t cap, size_t *len) {
  return guarded([&] {
    MSCX_REQUIRE(c && len, MSCX_EINVAL, "null argument");
    std::vector<uint32_t> d;
    c->c.dump(d);
    *len = d.size();
    if (buf && cap >= d.size()) memcpy(buf, d.data(), d.size() * 4);
  });
}

int mscx_complex_num_cps(const mscx_complex *c, uint32_t *n) {
  return guarded([&] {
    MSCX_REQUIRE(c && n, MSCX_EINVAL, "null argument");
    *n = c->c.size();
  });
}

int mscx_complex_destroy(mscx_complex *c) {
  return guarded([&] { delete c; });
}

}  // extern "C"
