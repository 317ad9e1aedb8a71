// The product's GridMSComplex algorithms (see complex.hpp).
#include "complex.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <memory>
#include <queue>
#include <sstream>
#include <stdexcept>

#include "runtime.hpp"
#include "tile.hpp"

namespace mscx {

// ------------------------------------------------------------------ CellMap

static inline uint32_t hash32(uint32_t k) {
  k ^= k >> 16;
  k *= 0x7FEB352Du;
  k ^= k >> 15;
  k *= 0x846CA68Bu;
  k ^= k >> 16;
  return k;
}

void CellMap::reserve(size_t n) {
  size_t want = 16;
  while (want < n * 2) want <<= 1;
  if (want <= keys_.size()) return;
  std::vector<uint32_t> ok, ov;
  ok.swap(keys_);
  ov.swap(vals_);
  keys_.assign(want, UINT32_MAX);
  vals_.assign(want, 0);
  mask_ = want - 1;
  n_ = 0;
  for (size_t i = 0; i < ok.size(); ++i)
    if (ok[i] != UINT32_MAX) set(ok[i], ov[i]);
}

void CellMap::grow() { reserve(keys_.empty() ? 16 : keys_.size()); }

size_t CellMap::slot(uint32_t key) const {
  size_t i = hash32(key) & mask_;
  while (keys_[i] != UINT32_MAX && keys_[i] != key) i = (i + 1) & mask_;
  return i;
}

uint32_t CellMap::find(uint32_t key) const {
  if (keys_.empty()) return UINT32_MAX;
  size_t i = slot(key);
  return keys_[i] == key ? vals_[i] : UINT32_MAX;
}

void CellMap::insert(uint32_t key, uint32_t val) {
  if ((n_ + 1) * 2 > keys_.size()) grow();
  size_t i = slot(key);
  if (keys_[i] == key) return;
  keys_[i] = key;
  vals_[i] = val;
  ++n_;
}

void CellMap::set(uint32_t key, uint32_t val) {
  if ((n_ + 1) * 2 > keys_.size()) grow();
  size_t i = slot(key);
  if (keys_[i] != key) {
    keys_[i] = key;
    ++n_;
  }
  vals_[i] = val;
}

uint32_t CellMap::get_or_zero(uint32_t key) {
  uint32_t v = find(key);
  if (v != UINT32_MAX) return v;
  set(key, 0);
  return 0;
}

// ------------------------------------------------------------------ Complex

uint32_t Complex::replicate(const Complex &src, uint32_t j) {
  uint32_t c = src.cell[j];
  if (idx.contains(c)) throw std::logic_error("this cp is present in msc");
  uint32_t n = size();
  cell.push_back(c);
  pair.push_back(UINT32_MAX);
  flags.push_back(src.flags[j]);
  asc.emplace_back();
  des.emplace_back();
  asc_disc.emplace_back();
  des_disc.emplace_back();
  idx.set(c, n);
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
    out.push_back((uint32_t)asc_disc[i].size());
    out.push_back((uint32_t)des_disc[i].size());
    out.insert(out.end(), asc[i].begin(), asc[i].end());
    out.insert(out.end(), des[i].begin(), des[i].end());
    out.insert(out.end(), asc_disc[i].begin(), asc_disc[i].end());
    out.insert(out.end(), des_disc[i].begin(), des_disc[i].end());
  }
}

// writeout_connectivity (grid_dataset.cpp:1858-1907) and
// addCriticalPointsToMSComplex (discreteMorseAlgorithm.h:244-268): nodes in
// raster order; each saddle connects to the owners of its facets (des) and
// cofacets (asc).  Arc multisets are order-free, so they are built in two
// passes (counts, then fills in index order, already sorted).
void complex_build(Complex &c, const Rect &rect, const Rect &ext, const std::vector<uint32_t> &crit,
                   const std::vector<float> &fn, const std::vector<uint32_t> &conn4,
                   const std::vector<uint32_t> &pair) {
  const uint32_t K = (uint32_t)crit.size();
  c.rect = rect;
  c.ext = ext;
  c.cell = crit;
  c.fn = fn;
  c.pair.assign(K, UINT32_MAX);
  c.flags.assign(K, 0);
  c.asc.assign(K, ConnSet());
  c.des.assign(K, ConnSet());
  c.asc_disc.assign(K, {});
  c.des_disc.assign(K, {});
  c.idx.reserve(K);
  for (uint32_t k = 0; k < K; ++k) c.idx.insert(crit[k], k);
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
  // extremum side: saddles visited in increasing index, so appends stay sorted
  for (uint32_t k = 0; k < K; ++k) {
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
  complex_build(out, r, e, crit, fn, conn, pr);
}

static constexpr uint64_t kMaxArcProduct = 1ull << 20;

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

static void merge_ixn(const Complex &a, const Complex &b, Rect &ixn) {
  // rect_t::intersection (rectangle_complex.h:208-220), called as
  // msc2.m_rect.intersection(msc1.m_rect)
  const Rect &r1 = b.rect, &r2 = a.rect;
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

// merge_up (grid_mscomplex.cpp:24-156)
void merge_up(const Complex &m1, const Complex &m2, Complex &res) {
  Rect ixn;
  merge_ixn(m1, m2, ixn);
  int bx, by, tx, ty;
  pmin(m1.rect.x0, m1.rect.y0, m2.rect.x0, m2.rect.y0, bx, by);
  pmax(m1.rect.x1, m1.rect.y1, m2.rect.x1, m2.rect.y1, tx, ty);
  res = Complex();
  Complex *out = &res;
  out->rect = rect_of_points(bx, by, tx, ty);
  pmin(m1.ext.x0, m1.ext.y0, m2.ext.x0, m2.ext.y0, bx, by);
  pmax(m1.ext.x1, m1.ext.y1, m2.ext.x1, m2.ext.y1, tx, ty);
  out->ext = rect_of_points(bx, by, tx, ty);
  const Complex *arr[2] = {&m1, &m2};
  size_t total = m1.size() + m2.size();
  out->idx.reserve(total);
  out->cell.reserve(total);
  out->fn.reserve(total);

  for (int i = 0; i < 2; ++i) {
    const Complex &m = *arr[i];
    for (uint32_t j = 0; j < m.size(); ++j) {
      if (i == 1 && out->idx.contains(m.cell[j])) continue;
      if (m.flags[j] & CPF_CANCELLED) continue;
      out->replicate(m, j);
      out->fn.push_back(m.fn[j]);
    }
  }
  for (int i = 0; i < 2; ++i) {
    const Complex &m = *arr[i];
    for (uint32_t j = 0; j < m.size(); ++j) {
      if (m.flags[j] & CPF_CANCELLED) continue;
      uint32_t di = out->idx.get_or_zero(m.cell[j]);
      if (m.flags[j] & CPF_BCANC) out->pair[di] = out->idx.get_or_zero(m.cell[m.pair[j]]);
      bool cmn = (i == 1) && ixn.contains(cx_of(m.cell[j]), cy_of(m.cell[j]));
      for (int k = 0; k < 2; ++k) {
        const ConnSet &src = k == 0 ? m.asc[j] : m.des[j];
        ConnSet &dst = k == 0 ? out->asc[di] : out->des[di];
        for (uint32_t v : src) {
          uint32_t vc = m.cell[v];
          if (cmn && ixn.contains(cx_of(vc), cy_of(vc))) continue;
          if (m.flags[v] & CPF_CANCELLED) continue;
          dst.insert(out->idx.get_or_zero(vc));
        }
      }
    }
  }
  for (int y = ixn.y0; y <= ixn.y1; ++y) {
    for (int x = ixn.x0; x <= ixn.x1; ++x) {
      uint32_t c = (uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16);
      uint32_t si = out->idx.find(c);
      if (si == UINT32_MAX) throw std::logic_error("missing common bndry cp");
      if ((out->flags[si] & CPF_CANCELLED) || !(out->flags[si] & CPF_BCANC)) continue;
      uint32_t pi = out->pair[si];
      uint32_t p = out->cell[pi];
      if (!out->rect.interior(x, y) && !out->ext.on_boundary(x, y)) continue;
      if (!out->rect.interior(cx_of(p), cy_of(p)) && !out->ext.on_boundary(cx_of(p), cy_of(p))) continue;
      cancel_pairs(*out, si, pi, nullptr);
    }
  }
}

// merge_down (grid_mscomplex.cpp:158-322)
void merge_down(Complex &P, Complex &m1, Complex &m2) {
  Rect ixn;
  merge_ixn(m1, m2, ixn);
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
  Complex *arr[2] = {&m1, &m2};
  for (int i = 0; i < 2; ++i) {
    Complex &m = *arr[i];
    for (uint32_t j = 0; j < P.size(); ++j) {
      if (P.flags[j] & CPF_CANCELLED) throw std::logic_error("all cps ought to be uncancelled by now");
      if (!(P.flags[j] & CPF_BCANC)) continue;
      uint32_t sp = P.pair[j];
      bool src_in = m.idx.contains(P.cell[j]);
      bool sp_in = m.idx.contains(P.cell[sp]);
      if (!src_in && !sp_in) continue;
      if (!src_in) {
        m.replicate(P, j);
        m.fn.push_back(P.fn[j]);
      }
      if (!sp_in) {
        m.replicate(P, sp);
        m.fn.push_back(P.fn[sp]);
      }
      uint32_t d = m.idx.find(P.cell[j]);
      if (!src_in || !sp_in || !(m.flags[d] & CPF_BCANC)) {
        uint32_t dp = m.idx.find(P.cell[sp]);
        m.flags[d] |= CPF_BCANC;
        m.flags[dp] |= CPF_BCANC;
        m.pair[d] = dp;
        m.pair[dp] = d;
      }
      for (int k = 0; k < 2; ++k) {
        const ConnSet &src = k == 0 ? P.asc[j] : P.des[j];
        // `dst` is re-fetched after replicate() may have grown the arrays
        (k == 0 ? m.asc[d] : m.des[d]).clear();
        for (uint32_t v : src) {
          if (P.flags[v] & CPF_BCANC) throw std::logic_error("only non cancellable cps must be remaining");
          uint32_t mv = m.idx.find(P.cell[v]);
          if (mv == UINT32_MAX) {
            mv = m.replicate(P, v);
            m.fn.push_back(P.fn[v]);
          }
          (k == 0 ? m.asc[d] : m.des[d]).insert(mv);
        }
      }
    }
    for (uint32_t j = 0; j < P.size(); ++j) {
      if (P.flags[j] & CPF_BCANC) continue;
      uint32_t d = m.idx.find(P.cell[j]);
      if (d == UINT32_MAX) continue;
      for (int k = 0; k < 2; ++k) {
        const ConnSet &src = k == 0 ? P.asc[j] : P.des[j];
        ConnSet &dst = k == 0 ? m.asc[d] : m.des[d];
        dst.clear();
        for (uint32_t v : src) {
          uint32_t mv = m.idx.find(P.cell[v]);
          if (mv == UINT32_MAX) continue;
          dst.insert(mv);
        }
      }
    }
  }
}

// persistence_comparator_t (grid_mscomplex.cpp:331-379).  Its geometric
// tie-break compares the first pair with itself, so two pairs of equal
// persistence are equivalent; only |f2 - f1| orders the heap.  The heap
// element carries that persistence (the function values do not change while
// simplifying), so a comparison reads no other memory; the sequence of heap
// operations, and hence the order among equal persistences, is the one
// std::priority_queue produces for the reference's (uint, uint) elements.
struct HeapPair {
  float pers;
  uint32_t first, second;
};
struct PersistenceLess {
  bool operator()(const HeapPair &p1, const HeapPair &p2) const {
    if (p1.pers != p2.pers) return p1.pers > p2.pers;
    return false;
  }
};

// simplify (grid_mscomplex.cpp:384-512)
void simplify(Complex &m, std::vector<std::pair<uint32_t, uint32_t>> &canc, double thresh) {
  typedef std::pair<uint32_t, uint32_t> P2;
  const float *fn = m.fn.data();
  auto mk = [fn](uint32_t a, uint32_t b) { return HeapPair{std::abs(fn[b] - fn[a]), a, b}; };
  std::vector<HeapPair> heap_store;
  size_t narcs = 0;
  for (uint32_t i = 0; i < m.size(); ++i) narcs += m.des[i].size();
  m.arc_budget = (int64_t)(16 * narcs + ((size_t)1 << 20));
  heap_store.reserve(narcs * 2 + 16);
  std::priority_queue<HeapPair, std::vector<HeapPair>, PersistenceLess> pq(PersistenceLess{}, std::move(heap_store));

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
  std::vector<HeapPair> resubmit;
  float max_persistence = max_val - min_val;
  std::vector<uint32_t> new_edges;
  const Rect &R = m.rect;
  auto on_b = [&](uint32_t c) { return R.on_boundary(cx_of(c), cy_of(c)); };

  while (!pq.empty()) {
    HeapPair hp = pq.top();
    pq.pop();
    uint32_t v1 = hp.first, v2 = hp.second;
    float pers = std::abs(m.fn[v1] - m.fn[v2]);
    if ((double)pers / (double)max_persistence > thresh) break;
    if (cell_dim(m.cell[v2]) == 1) std::swap(v1, v2);
    uint32_t d2 = cell_dim(m.cell[v2]);
    if ((m.flags[v1] & CPF_CANCELLED) || (m.flags[v2] & CPF_CANCELLED)) continue;
    bool b1 = on_b(m.cell[v1]), b2 = on_b(m.cell[v2]);
    if (b1 && !b2) continue;
    if (!b1 && b2) continue;
    const ConnSet &c2 = (d2 / 2 == 0) ? m.asc[v2] : m.des[v2];
    bool strang = m.flags[v1] & CPF_STRANGULATION;
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
    canc.push_back(P2(hp.first, hp.second));
    for (size_t i = 0; i < new_edges.size(); i += 2) pq.push(mk(new_edges[i], new_edges[i + 1]));
    for (const HeapPair &r : resubmit) pq.push(r);
    resubmit.clear();
  }
}

static int64_t arc_count(const Complex &m) {
  int64_t n = 0;
  for (uint32_t i = 0; i < m.size(); ++i) n += m.asc[i].size() + m.des[i].size();
  return n;
}

void un_simplify(Complex &m, const std::vector<std::pair<uint32_t, uint32_t>> &canc) {
  m.arc_budget = 16 * arc_count(m) + ((int64_t)1 << 20);
  for (auto it = canc.rbegin(); it != canc.rend(); ++it) uncancel_pairs(m, it->first, it->second);
  m.arc_budget = INT64_MAX;
}

// postMergeFillDiscs (grid_dataset.cpp:1044-1231).  codes / own are the
// tile's arrays over ext, x fastest; own is modified (owner redirects).
void fill_discs(Complex &m, const Rect &rect, const Rect &ext, const uint8_t *codes, std::vector<uint32_t> &own) {
  const int cw = ext.x1 - ext.x0 + 1;
  auto at = [&](int x, int y) { return (size_t)(x - ext.x0) + (size_t)(y - ext.y0) * cw; };
  auto code = [&](int x, int y) { return codes[at(x, y)]; };
  auto paired = [&](int x, int y) { return (code(x, y) & CF_PAIRED) != 0; };
  auto critical = [&](int x, int y) { return (code(x, y) & CF_CRITICAL) != 0; };
  auto pair_of = [&](int x, int y) -> uint32_t {
    int d = code_dir(code(x, y));
    if (!d) return UINT32_MAX;
    return (uint32_t)(uint16_t)(x + dir_dx(d)) | ((uint32_t)(uint16_t)(y + dir_dy(d)) << 16);
  };
  auto in_r = [&](uint32_t c) { return c != UINT32_MAX && rect.contains(cx_of(c), cy_of(c)); };

  CellMap asc_map, des_map;  // surviving extremum cell -> cp index
  std::vector<uint32_t> surv_saddles;
  for (uint32_t i = 0; i < m.size(); ++i) {
    uint32_t c = m.cell[i];
    uint32_t dim = cell_dim(c);
    if (m.flags[i] & CPF_BCANC) {
      uint32_t pi = m.pair[i];
      uint32_t pc = m.cell[pi];
      if (dim == 1) {
        const ConnSet &conn = (cell_dim(pc) == 0) ? m.des[i] : m.asc[i];
        if (conn.size() != 1) throw std::logic_error("I should be connected to exactly one surv extrema");
        uint32_t ec = m.cell[conn[0]];
        if (m.pair[pi] != i) throw std::logic_error("My pair does not know Im paired with him");
        if (in_r(c)) own[at(cx_of(c), cy_of(c))] = ec;
        if (in_r(pc)) own[at(cx_of(pc), cy_of(pc))] = ec;
      } else if (dim == 2) {
        for (uint32_t s : m.des[i])
          if (in_r(pc)) m.asc_disc[s].push_back(pc);
      } else {
        for (uint32_t s : m.asc[i])
          if (in_r(pc)) m.des_disc[s].push_back(pc);
      }
    } else {
      if (dim == 0) {
        asc_map.insert(c, i);
      } else if (dim == 1) {
        int x = cx_of(c), y = cy_of(c);
        const int inc[4][2] = {{x, y + 1}, {x, y - 1}, {x - 1, y}, {x + 1, y}};
        for (int j = 0; j < 4; ++j) {
          int ix = inc[j][0], iy = inc[j][1];
          if (rect.contains(ix, iy) && paired(ix, iy) && !critical(ix, iy)) {
            uint32_t ic = (uint32_t)(uint16_t)ix | ((uint32_t)(uint16_t)iy << 16);
            (cell_dim(ic) == 0 ? m.des_disc[i] : m.asc_disc[i]).push_back(pair_of(ix, iy));
          }
        }
        surv_saddles.push_back(i);
      } else {
        des_map.insert(c, i);
      }
    }
  }

  for (int y = rect.y0; y <= rect.y1; y += 2) {
    for (int x = rect.x0; x <= rect.x1; x += 2) {
      uint32_t o = own[at(x, y)];
      if (in_r(o)) o = own[at(cx_of(o), cy_of(o))];
      if (cx_of(o) != -1 && cy_of(o) != -1) {
        uint32_t k = asc_map.find(o);
        if (k == UINT32_MAX) throw std::logic_error("owner extremum not in surviving map");
        m.asc_disc[k].push_back((uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16));
      }
    }
  }
  for (int y = rect.y0 + 1; y < rect.y1; y += 2) {
    for (int x = rect.x0 + 1; x < rect.x1; x += 2) {
      uint32_t o = own[at(x, y)];
      if (in_r(o)) o = own[at(cx_of(o), cy_of(o))];
      if (cx_of(o) != -1 && cy_of(o) != -1) {
        uint32_t k = des_map.find(o);
        if (k == UINT32_MAX) throw std::logic_error("owner extremum not in surviving map");
        m.des_disc[k].push_back((uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16));
      }
    }
  }

  for (uint32_t si : surv_saddles) {
    for (int j = 0; j < 2; ++j) {
      std::vector<uint32_t> &disc = j == 0 ? m.asc_disc[si] : m.des_disc[si];
      for (size_t k = 0; k != disc.size(); ++k) {
        uint32_t pcell = disc[k];
        int px = cx_of(pcell), py = cy_of(pcell);
        uint32_t cets[4];
        int ct = 0;
        int dx = px & 1, dy = py & 1;
        if (j == 0) {  // cofacets, filtered to ext
          int cand[2][2];
          int nc = 0;
          if (dx + dy == 0) {
            const int c4[4][2] = {{px, py + 1}, {px, py - 1}, {px - 1, py}, {px + 1, py}};
            for (int q = 0; q < 4; ++q)
              if (ext.contains(c4[q][0], c4[q][1]))
                cets[ct++] = (uint32_t)(uint16_t)c4[q][0] | ((uint32_t)(uint16_t)c4[q][1] << 16);
          } else if (dx + dy == 1) {
            cand[0][0] = px + dy; cand[0][1] = py + dx;
            cand[1][0] = px - dy; cand[1][1] = py - dx;
            nc = 2;
            for (int q = 0; q < nc; ++q)
              if (ext.contains(cand[q][0], cand[q][1]))
                cets[ct++] = (uint32_t)(uint16_t)cand[q][0] | ((uint32_t)(uint16_t)cand[q][1] << 16);
          }
        } else {  // facets
          if (dx + dy == 1) {
            cets[ct++] = (uint32_t)(uint16_t)(px + dx) | ((uint32_t)(uint16_t)(py + dy) << 16);
            cets[ct++] = (uint32_t)(uint16_t)(px - dx) | ((uint32_t)(uint16_t)(py - dy) << 16);
          } else if (dx + dy == 2) {
            const int f4[4][2] = {{px, py + 1}, {px, py - 1}, {px - 1, py}, {px + 1, py}};
            for (int q = 0; q < 4; ++q)
              cets[ct++] = (uint32_t)(uint16_t)f4[q][0] | ((uint32_t)(uint16_t)f4[q][1] << 16);
          }
        }
        for (int q = 0; q < ct; ++q) {
          int qx = cx_of(cets[q]), qy = cy_of(cets[q]);
          if (critical(qx, qy)) continue;
          if (!rect.contains(qx, qy)) continue;
          uint32_t p = pair_of(qx, qy);
          if (p == pcell) continue;
          disc.push_back(p);
        }
      }
    }
  }
}

// write_discs (grid_mscomplex.cpp:542-590): raw packed ids appended to
// <prefix>asc_(x,y) and <prefix>des_(x,y).
void write_discs(const Complex &m, const std::string &prefix) {
  for (uint32_t i = 0; i < m.size(); ++i) {
    if (m.flags[i] & CPF_BCANC) continue;
    for (int j = 0; j < 2; ++j) {
      const std::vector<uint32_t> &d = j == 0 ? m.asc_disc[i] : m.des_disc[i];
      if (d.empty()) continue;
      std::ostringstream ss;
      ss << prefix << (j == 0 ? "asc_" : "des_") << "(" << cx_of(m.cell[i]) << "," << cy_of(m.cell[i]) << ")";
      std::ofstream os(ss.str(), std::ios::out | std::ios::ate | std::ios::app | std::ios::binary);
      if (!os.is_open()) throw Error(MSCX_EINVAL, "asc/des disc file not writeable");
      os.write((const char *)d.data(), (std::streamsize)(d.size() * 4));
    }
  }
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
    complex_build(h->c, r, e, std::vector<uint32_t>(crit, crit + K), std::vector<float>(fn, fn + K),
                  std::vector<uint32_t>(conn4, conn4 + 4 * (size_t)K), std::vector<uint32_t>(pair_idx, pair_idx + K));
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
    std::vector<uint8_t> codes;
    std::vector<uint32_t> own;
    tile_codes(t, codes);
    tile_owners(t, own);
    Rect r{t->rect[0], t->rect[1], t->rect[2], t->rect[3]}, e{t->ext[0], t->ext[1], t->ext[2], t->ext[3]};
    fill_discs(c->c, r, e, codes.data(), own);
  });
}

int mscx_complex_write_discs(const mscx_complex *c, const char *prefix) {
  return guarded([&] {
    MSCX_REQUIRE(c && prefix, MSCX_EINVAL, "null argument");
    write_discs(c->c, prefix);
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
