// oracle/oracle.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// CPU restatement of the reference grid path.  Every function names the
// reference lines it follows.  Deviations from the reference are limited to
// (a) defined behaviour where the reference has undefined behaviour (noted
// inline) and (b) the input window of a 2-D tile being copied into its own
// buffer instead of pointing into a shared slab buffer.
#include "oracle.hpp"

#include <algorithm>
#include <atomic>
#include <exception>
#include <functional>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <queue>
#include <stdexcept>

namespace orc {

// Planted mistakes for the pin tests (tests/test_oracle_pins_cpu.py): 0 in
// every normal run.  Each value makes one plausible error in one function,
// so that a test can show that its check fails on that error.
//   1 merge_up keeps the shared line's duplicate arcs (no dedup)
//   2 cancelPairs adds new arcs with set semantics (drops multiplicity)
//   3 uncancel_pairs keeps only non-cancellable neighbours (no expansion
//     through a cancelled neighbour's pair)
//   4 simplify starts max_val at the lowest float instead of FLT_MIN
int g_plant = 0;
// Last phase DataManager::run performs (0 compute .. 4 discs); the pin tests
// lower it to look at a planted mistake before a later phase's own
// invariant checks throw.
int g_stop_after = 4;

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------- GridDataset

// grid_dataset.cpp:1461-1484.  Flags are zeroed, pairs/owners default to
// (-1,-1) as boost::multi_array default-constructs cellid_t.
void GridDataset::init(const fn_t *data) {
  cellid_t s = m_ext_rect.size();
  if (data) m_verts = data;
  m_vx = 1 + s[0] / 2;
  m_vy = 1 + s[1] / 2;
  m_cx = 1 + s[0];
  m_cy = 1 + s[1];
  size_t n = (size_t)m_cx * m_cy;
  m_cell_flags.assign(n, CELLFLAG_UNKNOWN);
  m_cell_pairs.assign(n, cellid_t());
  m_cell_own.assign(n, cellid_t());
  // m_critical_cells is deliberately not cleared here (grid_dataset.cpp:1461
  // does not clear it; a second work() appends a duplicate list).
}

// grid_dataset.cpp:1486-1509
void GridDataset::clear() {
  m_cell_flags.clear();
  m_cell_pairs.clear();
  m_cell_own.clear();
  m_critical_cells.clear();
}

// grid_dataset.h:158-167: value compare (as double), ties broken by cellid
// order.
bool GridDataset::ptLt(cellid_t c1, cellid_t c2) const {
  double f1 = vert(c1);
  double f2 = vert(c2);
  if (f1 != f2) return f1 < f2;
  return c1 < c2;
}

// grid_dataset.cpp:1566-1591
uint GridDataset::getCellPoints(cellid_t c, cellid_t *p) const {
  switch (getCellDim(c)) {
    case 0:
      p[0] = c;
      return 1;
    case 1: {
      int d0 = c[0] & 1, d1 = c[1] & 1;
      p[0] = cellid_t(c[0] + d0, c[1] + d1);
      p[1] = cellid_t(c[0] - d0, c[1] - d1);
      return 2;
    }
    case 2:
      p[0] = cellid_t(c[0] + 1, c[1] + 1);
      p[1] = cellid_t(c[0] + 1, c[1] - 1);
      p[2] = cellid_t(c[0] - 1, c[1] - 1);
      p[3] = cellid_t(c[0] - 1, c[1] + 1);
      return 4;
  }
  throw std::logic_error("impossible dim");
}

// grid_dataset.cpp:1593-1617 (facets are not filtered to the ext rect)
uint GridDataset::getCellFacets(cellid_t c, cellid_t *f) const {
  switch (getCellDim(c)) {
    case 0:
      return 0;
    case 1: {
      int d0 = c[0] & 1, d1 = c[1] & 1;
      f[0] = cellid_t(c[0] + d0, c[1] + d1);
      f[1] = cellid_t(c[0] - d0, c[1] - d1);
      return 2;
    }
    case 2:
      f[0] = cellid_t(c[0], c[1] + 1);
      f[1] = cellid_t(c[0], c[1] - 1);
      f[2] = cellid_t(c[0] - 1, c[1]);
      f[3] = cellid_t(c[0] + 1, c[1]);
      return 4;
  }
  throw std::logic_error("impossible dim");
}

// grid_dataset.cpp:1619-1657 (cofacets are filtered to the ext rect)
uint GridDataset::getCellCofacets(cellid_t c, cellid_t *cf) const {
  uint ct = 0;
  switch (getCellDim(c)) {
    case 0:
      cf[0] = cellid_t(c[0], c[1] + 1);
      cf[1] = cellid_t(c[0], c[1] - 1);
      cf[2] = cellid_t(c[0] - 1, c[1]);
      cf[3] = cellid_t(c[0] + 1, c[1]);
      ct = 4;
      break;
    case 1: {
      int d0 = c[0] & 1, d1 = c[1] & 1;
      cf[0] = cellid_t(c[0] + d1, c[1] + d0);
      cf[1] = cellid_t(c[0] - d1, c[1] - d0);
      ct = 2;
      break;
    }
    case 2:
      return 0;
  }
  uint pos = 0;
  for (uint i = 0; i < ct; ++i)
    if (m_ext_rect.contains(cf[i])) cf[pos++] = cf[i];
  return pos;
}

// grid_dataset.cpp:1035-1042
uint GridDataset::getCellIncCells(cellid_t c, cellid_t *inc) const {
  inc[0] = cellid_t(c[0], c[1] + 1);
  inc[1] = cellid_t(c[0], c[1] - 1);
  inc[2] = cellid_t(c[0] - 1, c[1]);
  inc[3] = cellid_t(c[0] + 1, c[1]);
  return 4;
}

// grid_dataset.cpp:1519-1536: points sorted ascending, then compared
// lexicographically.
bool GridDataset::compareCells(cellid_t c1, cellid_t c2) const {
  if (getCellDim(c1) == 0) return ptLt(c1, c2);
  cellid_t p1[4], p2[4];
  uint n1 = getCellPoints(c1, p1), n2 = getCellPoints(c2, p2);
  auto lt = [this](const cellid_t &a, const cellid_t &b) { return ptLt(a, b); };
  std::sort(p1, p1 + n1, lt);
  std::sort(p2, p2 + n2, lt);
  return std::lexicographical_compare(p1, p1 + n1, p2, p2 + n2, lt);
}

// grid_dataset.cpp:1538-1552: float accumulation in getCellPoints order.
fn_t GridDataset::get_cell_fn(cellid_t c) const {
  fn_t fn = 0.0;
  cellid_t pts[4];
  uint n = getCellPoints(c, pts);
  for (uint j = 0; j < n; ++j) fn += m_verts[(size_t)(pts[j][0] / 2 - m_ext_rect.bl[0] / 2) +
                                             (size_t)(pts[j][1] / 2 - m_ext_rect.bl[1] / 2) * m_vx];
  fn /= n;
  return fn;
}

// grid_dataset.cpp:1684-1691
void GridDataset::pairCells(cellid_t c, cellid_t p) {
  m_cell_pairs[cidx(c)] = p;
  m_cell_pairs[cidx(p)] = c;
  m_cell_flags[cidx(c)] |= CELLFLAG_PAIRED;
  m_cell_flags[cidx(p)] |= CELLFLAG_PAIRED;
}

// grid_dataset.cpp:1314-1373
static bool lowestPairableCoFacet(const GridDataset *ds, cellid_t c, cellid_t &pairid) {
  cellid_t cofacets[4];
  bool usable[4];
  uint cf_ct = ds->getCellCofacets(c, cofacets);
  bool c_tb = ds->isTrueBoundryCell(c);
  for (uint i = 0; i < cf_ct; i++) {
    cellid_t facets[4];
    uint f_ct = ds->getCellFacets(cofacets[i], facets);
    usable[i] = true;
    if (c_tb && !ds->isTrueBoundryCell(cofacets[i])) {
      usable[i] = false;
      continue;
    }
    for (uint j = 0; j < f_ct; j++) {
      if (ds->compareCells(c, facets[j])) {
        usable[i] = false;
        break;
      }
    }
  }
  bool found = false;
  for (uint i = 0; i < cf_ct; i++) {
    if (!usable[i]) continue;
    if (!found) {
      found = true;
      pairid = cofacets[i];
      continue;
    }
    if (ds->compareCells(cofacets[i], pairid)) pairid = cofacets[i];
  }
  return found;
}

// grid_dataset.cpp:1743-1820
void GridDataset::assignGradients() {
  for (int y = m_rect.bottom(); y <= m_rect.top(); ++y)
    for (int x = m_rect.left(); x <= m_rect.right(); ++x) {
      cellid_t c(x, y), p;
      if (isCellMarked(c)) continue;
      if (lowestPairableCoFacet(this, c, p)) pairCells(c, p);
    }

  for (int y = m_rect.bottom(); y <= m_rect.top(); ++y)
    for (int x = m_rect.left(); x <= m_rect.right(); ++x) {
      cellid_t c(x, y);
      if (!isCellMarked(c)) markCellCritical(c);
    }

  auto mark_bndry = [this](cellid_t c) {
    if (isCellCritical(c)) return;
    cellid_t cf[4];
    uint cf_ct = getCellCofacets(c, cf);
    for (uint j = 0; j < cf_ct; ++j) {
      if (isCellExterior(cf[j])) {
        markCellCritical(c);
        markCellCritical(getCellPairId(c));
        break;
      }
    }
  };
  for (int x = m_rect.left(); x <= m_rect.right(); ++x) {
    mark_bndry(cellid_t(x, m_rect.bottom()));
    mark_bndry(cellid_t(x, m_rect.top()));
  }
  for (int y = m_rect.bottom() + 1; y < m_rect.top(); ++y) {
    mark_bndry(cellid_t(m_rect.left(), y));
    mark_bndry(cellid_t(m_rect.right(), y));
  }
}

// grid_dataset.cpp:1822-1832 (y-major raster over the ext rect)
void GridDataset::collateCriticalPoints() {
  for (int y = m_ext_rect.bottom(); y <= m_ext_rect.top(); ++y)
    for (int x = m_ext_rect.left(); x <= m_ext_rect.right(); ++x) {
      cellid_t c(x, y);
      if (isCellCritical(c)) m_critical_cells.push_back(c);
    }
}

// grid_dataset.cpp:1376-1427.  dir 0 = facets (downward), 1 = cofacets.
static void track_gradient_tree_bfs(GridDataset *ds, cellid_t start, int dir) {
  std::queue<cellid_t> q;
  q.push(start);
  while (!q.empty()) {
    cellid_t top = q.front();
    q.pop();
    ds->m_cell_own[ds->cidx(top)] = start;
    cellid_t cets[4];
    uint ct = dir ? ds->getCellCofacets(top, cets) : ds->getCellFacets(top, cets);
    for (uint i = 0; i < ct; i++) {
      if (ds->isCellCritical(cets[i])) continue;
      if (ds->isCellExterior(cets[i])) continue;
      cellid_t next = ds->getCellPairId(cets[i]);
      if (GridDataset::getCellDim(top) == GridDataset::getCellDim(next) && next != top) {
        ds->m_cell_own[ds->cidx(cets[i])] = start;
        q.push(next);
      }
    }
  }
}

// grid_dataset.cpp:1835-1856
void GridDataset::assignCellOwnerExtrema() {
  for (size_t k = 0; k < m_critical_cells.size(); ++k) {
    cellid_t c = m_critical_cells[k];
    m_cell_own[cidx(c)] = c;
    switch (getCellDim(c)) {
      case 0:
        track_gradient_tree_bfs(this, c, 1);
        break;
      case 2:
        track_gradient_tree_bfs(this, c, 0);
        break;
      default:
        break;
    }
  }
}

// grid_dataset.cpp:1734-1741
void GridDataset::work() {
  assignGradients();
  collateCriticalPoints();
  assignCellOwnerExtrema();
}

// grid_dataset.cpp:1285-1312
static void connectCps(GridMSComplex *msc, cellid_t c1, cellid_t c2) {
  if (GridDataset::getCellDim(c1) < GridDataset::getCellDim(c2)) std::swap(c1, c2);
  if (GridDataset::getCellDim(c1) != GridDataset::getCellDim(c2) + 1)
    throw std::logic_error("must connect i,i+1 cp (or vice versa)");
  auto i1 = msc->m_id_cp_map.find(c1), i2 = msc->m_id_cp_map.find(c2);
  if (i1 == msc->m_id_cp_map.end()) throw std::logic_error("cell not in id_cp_map c1");
  if (i2 == msc->m_id_cp_map.end()) throw std::logic_error("cell not in id_cp_map c2");
  uint a = i1->second, b = i2->second;
  msc->m_cps[a]->des.insert(b);
  msc->m_cps[b]->asc.insert(a);
}

// discreteMorseAlgorithm.h:244-268 (map::insert keeps the first index of a
// repeated cellid).
template <typename It>
static void addCriticalPointsToMSComplex(GridMSComplex *msc, It begin, It end) {
  msc->m_cps.resize(end - begin);
  uint pos = 0;
  for (It it = begin; it != end; ++it) {
    critpt_t *cp = new critpt_t;
    msc->m_cps[pos] = cp;
    cp->cellid = *it;
    msc->m_id_cp_map.insert(std::make_pair(*it, pos));
    ++pos;
  }
}

// grid_dataset.cpp:1858-1907
void GridDataset::writeout_connectivity(GridMSComplex *msgraph) {
  addCriticalPointsToMSComplex(msgraph, m_critical_cells.begin(), m_critical_cells.end());
  msgraph->m_cp_fns.resize(m_critical_cells.size());
  for (size_t k = 0; k < m_critical_cells.size(); ++k) {
    cellid_t c = m_critical_cells[k];
    uint cp_idx = msgraph->m_id_cp_map[c];
    msgraph->m_cp_fns[cp_idx] = get_cell_fn(c);
    if (getCellDim(c) == 1) {
      cellid_t f[4], cf[4];
      uint f_ct = getCellFacets(c, f), cf_ct = getCellCofacets(c, cf);
      for (uint i = 0; i < f_ct; ++i) {
        cellid_t o = m_cell_own[cidx(f[i])];
        if (o != cellid_t(-1, -1)) connectCps(msgraph, c, o);
      }
      for (uint i = 0; i < cf_ct; ++i) {
        cellid_t o = m_cell_own[cidx(cf[i])];
        if (o != cellid_t(-1, -1)) connectCps(msgraph, c, o);
      }
    }
    if (!isCellPaired(c)) continue;
    msgraph->m_cps[cp_idx]->isBoundryCancelable = true;
    msgraph->m_cps[cp_idx]->pair_idx = msgraph->m_id_cp_map[getCellPairId(c)];
  }
}

// grid_dataset.cpp:1044-1231
int GridDataset::postMergeFillDiscs(GridMSComplex *msgraph) {
  std::map<cellid_t, critpt_t::disc_t *> asc_map, des_map;
  std::vector<uint> surv_saddles;

  for (uint i = 0; i < msgraph->m_cps.size(); ++i) {
    critpt_t *cp = msgraph->m_cps[i];
    int dim = getCellDim(cp->cellid);
    if (cp->isBoundryCancelable) {
      critpt_t *pair_cp = msgraph->m_cps[cp->pair_idx];
      if (dim == 1) {
        critpt_t::conn_t *conn = (getCellDim(pair_cp->cellid) == 0) ? &cp->des : &cp->asc;
        if (conn->size() != 1) throw std::logic_error("I should be connected to exactly one surv extrema");
        critpt_t *ext_cp = msgraph->m_cps[*conn->begin()];
        if (pair_cp->pair_idx != i) throw std::logic_error("My pair does not know Im paired with him");
        if (m_rect.contains(cp->cellid)) m_cell_own[cidx(cp->cellid)] = ext_cp->cellid;
        if (m_rect.contains(pair_cp->cellid)) m_cell_own[cidx(pair_cp->cellid)] = ext_cp->cellid;
      } else if (dim == 2) {
        for (auto it = cp->des.begin(); it != cp->des.end(); ++it) {
          critpt_t *s = msgraph->m_cps[*it];
          if (m_rect.contains(pair_cp->cellid)) s->asc_disc.push_back(pair_cp->cellid);
        }
      } else if (dim == 0) {
        for (auto it = cp->asc.begin(); it != cp->asc.end(); ++it) {
          critpt_t *s = msgraph->m_cps[*it];
          if (m_rect.contains(pair_cp->cellid)) s->des_disc.push_back(pair_cp->cellid);
        }
      }
    } else {
      switch (dim) {
        case 0:
          asc_map.insert(std::make_pair(cp->cellid, &cp->asc_disc));
          break;
        case 1: {
          cellid_t inc[4];
          getCellIncCells(cp->cellid, inc);
          for (uint j = 0; j < 4; ++j) {
            if (m_rect.contains(inc[j]) && isCellPaired(inc[j]) && !isCellCritical(inc[j])) {
              critpt_t::disc_t *disc = (getCellDim(inc[j]) == 0) ? &cp->des_disc : &cp->asc_disc;
              disc->push_back(getCellPairId(inc[j]));
            }
          }
          surv_saddles.push_back(i);
          break;
        }
        case 2:
          des_map.insert(std::make_pair(cp->cellid, &cp->des_disc));
          break;
      }
    }
  }

  // The reference dereferences map::operator[] unconditionally; a missing
  // owner would be a null dereference there.  Here it is a logic_error.
  auto lookup = [](std::map<cellid_t, critpt_t::disc_t *> &m, cellid_t o) {
    auto it = m.find(o);
    if (it == m.end()) throw std::logic_error("owner extremum not in surviving map");
    return it->second;
  };

  for (int y = m_rect.bottom();; y += 2) {
    for (int x = m_rect.left();; x += 2) {
      cellid_t c(x, y);
      cellid_t o = m_cell_own[cidx(c)];
      if (m_rect.contains(o)) o = m_cell_own[cidx(o)];
      if (o[0] != -1 && o[1] != -1) lookup(asc_map, o)->push_back(c);
      if (x == m_rect.right()) break;
    }
    if (y == m_rect.top()) break;
  }
  for (int y = m_rect.bottom() + 1;; y += 2) {
    for (int x = m_rect.left() + 1;; x += 2) {
      cellid_t c(x, y);
      cellid_t o = m_cell_own[cidx(c)];
      if (m_rect.contains(o)) o = m_cell_own[cidx(o)];
      if (o[0] != -1 && o[1] != -1) lookup(des_map, o)->push_back(c);
      if (x + 1 == m_rect.right()) break;
    }
    if (y + 1 == m_rect.top()) break;
  }

  for (size_t i = 0; i < surv_saddles.size(); ++i) {
    critpt_t *cp = msgraph->m_cps[surv_saddles[i]];
    critpt_t::disc_t *disc[] = {&cp->asc_disc, &cp->des_disc};
    for (uint j = 0; j < 2; ++j) {
      size_t k = 0;
      while (k != disc[j]->size()) {
        cellid_t path_cell = (*disc[j])[k];
        cellid_t cets[4];
        // getcets[(j+1)%2]: j=0 -> cofacets, j=1 -> facets
        uint ct = (j == 0) ? getCellCofacets(path_cell, cets) : getCellFacets(path_cell, cets);
        for (uint m = 0; m < ct; ++m) {
          if (isCellCritical(cets[m])) continue;
          if (!m_rect.contains(cets[m])) continue;
          cellid_t p = getCellPairId(cets[m]);
          if (p == path_cell) continue;
          disc[j]->push_back(p);
        }
        k++;
      }
    }
  }
  return 0;
}

// ------------------------------------------------------------------ MSComplex

void GridMSComplex::clear() {
  for (critpt_t *cp : m_cps) delete cp;
  m_cps.clear();
  m_id_cp_map.clear();
}

void GridMSComplex::dump(std::vector<uint32_t> &out) const {
  out.clear();
  out.push_back((uint32_t)m_cps.size());
  const rect_t *rs[2] = {&m_rect, &m_ext_rect};
  for (const rect_t *r : rs) {
    out.push_back((uint32_t)r->bl[0]);
    out.push_back((uint32_t)r->bl[1]);
    out.push_back((uint32_t)r->tr[0]);
    out.push_back((uint32_t)r->tr[1]);
  }
  for (size_t i = 0; i < m_cps.size(); ++i) {
    const critpt_t *cp = m_cps[i];
    uint32_t fb;
    fn_t f = i < m_cp_fns.size() ? m_cp_fns[i] : 0.0f;
    memcpy(&fb, &f, 4);
    out.push_back(cp->cellid.packed());
    out.push_back(fb);
    out.push_back(cp->pair_idx);
    out.push_back((cp->isCancelled ? 1u : 0u) | (cp->isOnStrangulationPath ? 2u : 0u) |
                  (cp->isBoundryCancelable ? 4u : 0u));
    out.push_back((uint32_t)cp->asc.size());
    out.push_back((uint32_t)cp->des.size());
    out.push_back((uint32_t)cp->asc_disc.size());
    out.push_back((uint32_t)cp->des_disc.size());
    for (uint a : cp->asc) out.push_back(a);
    for (uint d : cp->des) out.push_back(d);
    for (const cellid_t &c : cp->asc_disc) out.push_back(c.packed());
    for (const cellid_t &c : cp->des_disc) out.push_back(c.packed());
  }
}

// discreteMorseAlgorithm.h:348-464 and :466-585 (the variant that records
// new edges).  cp0 becomes the lower-dimensional cell (its asc holds cp1).
void cancelPairs(GridMSComplex *msc, uint cp0, uint cp1, std::vector<uint> *new_edges) {
  if (msc->m_cps[cp0]->asc.find(cp1) == msc->m_cps[cp0]->asc.end()) std::swap(cp0, cp1);
  if (msc->m_cps[cp0]->asc.find(cp1) == msc->m_cps[cp0]->asc.end()) return;  // logged error only

  critpt_t *p0 = msc->m_cps[cp0], *p1 = msc->m_cps[cp1];
  for (auto a = p0->asc.begin(); a != p0->asc.end(); ++a) {
    for (auto d = p1->des.begin(); d != p1->des.end(); ++d) {
      if (*a == cp1 || *d == cp0) continue;
      if (g_plant == 2 && msc->m_cps[*a]->des.count(*d)) continue;
      msc->m_cps[*a]->des.insert(*d);
      msc->m_cps[*d]->asc.insert(*a);
      if (new_edges) {
        new_edges->push_back(*d);
        new_edges->push_back(*a);
      }
      if (msc->m_cps[*a]->des.count(*d) >= 2) {
        msc->m_cps[*a]->isOnStrangulationPath = true;
        msc->m_cps[*d]->isOnStrangulationPath = true;
      }
    }
  }
  for (auto a = p0->asc.begin(); a != p0->asc.end(); ++a) {
    if (*a == cp1) continue;
    msc->m_cps[*a]->des.erase(cp0);
  }
  for (auto d = p0->des.begin(); d != p0->des.end(); ++d) msc->m_cps[*d]->asc.erase(cp0);
  for (auto a = p1->asc.begin(); a != p1->asc.end(); ++a) msc->m_cps[*a]->des.erase(cp1);
  for (auto d = p1->des.begin(); d != p1->des.end(); ++d) {
    if (*d == cp0) continue;
    msc->m_cps[*d]->asc.erase(cp1);
  }
  p0->isCancelled = true;
  p0->des.clear();
  p1->isCancelled = true;
  p1->asc.clear();
}

// discreteMorseAlgorithm.h:1006-1089
void uncancel_pairs(GridMSComplex *msc, uint cp1, uint cp2) {
  if (msc->m_cps[cp1]->asc.find(cp2) == msc->m_cps[cp1]->asc.end()) std::swap(cp1, cp2);
  critpt_t::conn_t *acdc[] = {&msc->m_cps[cp1]->asc, &msc->m_cps[cp2]->des};
  uint idxs[] = {cp1, cp2};
  for (uint i = 0; i < 2; ++i) {
    uint other = idxs[(i + 1) % 2];
    if (acdc[i]->find(other) == acdc[i]->end()) throw std::logic_error("cancellable pair is not connected");
    critpt_t::conn_t fresh;
    for (auto it = acdc[i]->begin(); it != acdc[i]->end(); ++it) {
      if (*it == other) continue;
      critpt_t *conn = msc->m_cps[*it];
      if (!conn->isBoundryCancelable || g_plant == 3) {
        fresh.insert(*it);
        continue;
      }
      if (conn->isCancelled) throw std::logic_error("*acdc_conn_it should not have been cancelled yet");
      uint pr = conn->pair_idx;
      if (msc->m_cps[pr]->pair_idx != *it) throw std::logic_error("*acdc_conn_it and its pair dont agree on pairing");
      critpt_t::conn_t *pc[] = {&msc->m_cps[pr]->asc, &msc->m_cps[pr]->des};
      for (auto jt = pc[i]->begin(); jt != pc[i]->end(); ++jt) {
        if (msc->m_cps[*jt]->isBoundryCancelable)
          throw std::logic_error("the connected cancelable cp shold not  contain a cancelable cp in its connections");
        fresh.insert(*jt);
      }
    }
    acdc[i]->clear();
    acdc[i]->insert(fresh.begin(), fresh.end());
  }
  msc->m_cps[cp1]->isCancelled = false;
  msc->m_cps[cp2]->isCancelled = false;
}

// grid_mscomplex.cpp:7-22
static void shallow_replicate_cp(GridMSComplex &msc, const critpt_t &cp) {
  if (msc.m_id_cp_map.count(cp.cellid) != 0) throw std::logic_error("this cp is present in msc");
  critpt_t *d = new critpt_t;
  d->isCancelled = cp.isCancelled;
  d->isBoundryCancelable = cp.isBoundryCancelable;
  d->isOnStrangulationPath = cp.isOnStrangulationPath;
  d->cellid = cp.cellid;
  msc.m_id_cp_map[d->cellid] = (uint)msc.m_cps.size();
  msc.m_cps.push_back(d);
}

static void check_merge_ixn(const GridMSComplex &a, const GridMSComplex &b, rect_t &ixn) {
  if (!b.m_rect.intersection(a.m_rect, ixn)) throw std::logic_error("rects should intersect for merge");
  if (ixn.left() != ixn.right() && ixn.top() != ixn.bottom())
    throw std::logic_error("rects must merge along a 1 manifold");
  if (ixn.bl == ixn.tr) throw std::logic_error("rects cannot merge along a point alone");
}

// grid_mscomplex.cpp:24-156
GridMSComplex *GridMSComplex::merge_up(const GridMSComplex &msc1, const GridMSComplex &msc2) {
  rect_t ixn;
  check_merge_ixn(msc1, msc2, ixn);
  rect_t r(std::min(msc1.m_rect.bl, msc2.m_rect.bl), std::max(msc1.m_rect.tr, msc2.m_rect.tr));
  rect_t e(std::min(msc1.m_ext_rect.bl, msc2.m_ext_rect.bl), std::max(msc1.m_ext_rect.tr, msc2.m_ext_rect.tr));
  GridMSComplex *out = new GridMSComplex(r, e);
  const GridMSComplex *arr[] = {&msc1, &msc2};

  for (uint i = 0; i < 2; ++i) {
    const GridMSComplex *msc = arr[i];
    for (uint j = 0; j < msc->m_cps.size(); ++j) {
      const critpt_t *src = msc->m_cps[j];
      if (i == 1 && out->m_id_cp_map.count(src->cellid) == 1) continue;
      if (src->isCancelled) continue;
      shallow_replicate_cp(*out, *src);
      out->m_cp_fns.push_back(msc->m_cp_fns[j]);
    }
  }

  for (uint i = 0; i < 2; ++i) {
    const GridMSComplex *msc = arr[i];
    for (uint j = 0; j < msc->m_cps.size(); ++j) {
      const critpt_t *src = msc->m_cps[j];
      if (src->isCancelled) continue;
      critpt_t *dst = out->m_cps[out->m_id_cp_map[src->cellid]];
      if (src->isBoundryCancelable) {
        const critpt_t *sp = msc->m_cps[src->pair_idx];
        dst->pair_idx = out->m_id_cp_map[sp->cellid];
      }
      const critpt_t::conn_t *s_acdc[] = {&src->asc, &src->des};
      critpt_t::conn_t *d_acdc[] = {&dst->asc, &dst->des};
      bool cmn = ixn.contains(src->cellid) && i == 1;
      for (uint k = 0; k < 2; ++k) {
        for (auto it = s_acdc[k]->begin(); it != s_acdc[k]->end(); ++it) {
          const critpt_t *conn = msc->m_cps[*it];
          if (cmn && ixn.contains(conn->cellid) && g_plant != 1) continue;
          if (conn->isCancelled) continue;
          d_acdc[k]->insert(out->m_id_cp_map[conn->cellid]);
        }
      }
    }
  }

  for (int y = ixn.bottom(); y <= ixn.top(); ++y) {
    for (int x = ixn.left(); x <= ixn.right(); ++x) {
      cellid_t c(x, y);
      auto f = out->m_id_cp_map.find(c);
      if (f == out->m_id_cp_map.end()) throw std::logic_error("missing common bndry cp");
      uint src_idx = f->second;
      critpt_t *src = out->m_cps[src_idx];
      if (src->isCancelled || !src->isBoundryCancelable) continue;
      uint pair_idx = src->pair_idx;
      cellid_t p = out->m_cps[pair_idx]->cellid;
      if (!out->m_rect.isInInterior(c) && !out->m_ext_rect.isOnBoundry(c)) continue;
      if (!out->m_rect.isInInterior(p) && !out->m_ext_rect.isOnBoundry(p)) continue;
      cancelPairs(out, src_idx, pair_idx, nullptr);
    }
  }
  return out;
}

// grid_mscomplex.cpp:158-322
void GridMSComplex::merge_down(GridMSComplex &msc1, GridMSComplex &msc2) {
  rect_t ixn;
  check_merge_ixn(msc1, msc2, ixn);

  for (int y = ixn.top(); y >= ixn.bottom(); --y) {
    for (int x = ixn.right(); x >= ixn.left(); --x) {
      cellid_t c(x, y);
      auto f = m_id_cp_map.find(c);
      if (f == m_id_cp_map.end()) throw std::logic_error("missing common bndry cp");
      uint src_idx = f->second;
      critpt_t *src = m_cps[src_idx];
      if (!src->isCancelled) continue;
      uint pair_idx = src->pair_idx;
      cellid_t p = m_cps[pair_idx]->cellid;
      if (!m_rect.isInInterior(c) && !m_ext_rect.isOnBoundry(c)) continue;
      if (!m_rect.isInInterior(p) && !m_ext_rect.isOnBoundry(p)) continue;
      uncancel_pairs(this, src_idx, pair_idx);
    }
  }

  GridMSComplex *arr[] = {&msc1, &msc2};
  for (uint i = 0; i < 2; ++i) {
    GridMSComplex *msc = arr[i];
    for (uint j = 0; j < m_cps.size(); ++j) {
      critpt_t *src = m_cps[j];
      if (src->isCancelled) throw std::logic_error("all cps ought to be uncancelled by now");
      if (!src->isBoundryCancelable) continue;
      critpt_t *sp = m_cps[src->pair_idx];
      bool src_in = msc->m_id_cp_map.count(src->cellid) != 0;
      bool sp_in = msc->m_id_cp_map.count(sp->cellid) != 0;
      if (!src_in && !sp_in) continue;
      if (!src_in) {
        shallow_replicate_cp(*msc, *src);
        msc->m_cp_fns.push_back(m_cp_fns[j]);
      }
      if (!sp_in) {
        shallow_replicate_cp(*msc, *sp);
        msc->m_cp_fns.push_back(m_cp_fns[src->pair_idx]);
      }
      uint d_idx = msc->m_id_cp_map[src->cellid];
      critpt_t *dst = msc->m_cps[d_idx];
      if (!src_in || !sp_in || !dst->isBoundryCancelable) {
        uint dp_idx = msc->m_id_cp_map[sp->cellid];
        critpt_t *dp = msc->m_cps[dp_idx];
        dst->isBoundryCancelable = true;
        dp->isBoundryCancelable = true;
        dst->pair_idx = dp_idx;
        dp->pair_idx = d_idx;
      }
      critpt_t::conn_t *s_acdc[] = {&src->asc, &src->des};
      critpt_t::conn_t *d_acdc[] = {&dst->asc, &dst->des};
      for (uint k = 0; k < 2; ++k) {
        d_acdc[k]->clear();
        for (auto it = s_acdc[k]->begin(); it != s_acdc[k]->end(); ++it) {
          critpt_t *sc = m_cps[*it];
          if (sc->isBoundryCancelable) throw std::logic_error("only non cancellable cps must be remaining");
          if (msc->m_id_cp_map.count(sc->cellid) == 0) {
            shallow_replicate_cp(*msc, *sc);
            msc->m_cp_fns.push_back(m_cp_fns[*it]);
          }
          d_acdc[k]->insert(msc->m_id_cp_map[sc->cellid]);
        }
      }
    }
    for (uint j = 0; j < m_cps.size(); ++j) {
      critpt_t *src = m_cps[j];
      if (src->isBoundryCancelable) continue;
      if (msc->m_id_cp_map.count(src->cellid) != 1) continue;
      critpt_t *dst = msc->m_cps[msc->m_id_cp_map[src->cellid]];
      critpt_t::conn_t *s_acdc[] = {&src->asc, &src->des};
      critpt_t::conn_t *d_acdc[] = {&dst->asc, &dst->des};
      for (uint k = 0; k < 2; ++k) {
        d_acdc[k]->clear();
        for (auto it = s_acdc[k]->begin(); it != s_acdc[k]->end(); ++it) {
          critpt_t *sc = m_cps[*it];
          auto f = msc->m_id_cp_map.find(sc->cellid);
          if (f == msc->m_id_cp_map.end()) continue;
          d_acdc[k]->insert(f->second);
        }
      }
    }
  }
}

// grid_mscomplex.cpp:331-379.  The geometric tie-break compares the first
// pair with itself (reference lines 355-359), so equal persistences compare
// as equivalent and their order is whatever the heap yields.  Restated as is.
struct persistence_comparator_t {
  GridMSComplex *m_msc;
  explicit persistence_comparator_t(GridMSComplex *m) : m_msc(m) {}
  bool operator()(const std::pair<uint, uint> &p1, const std::pair<uint, uint> &p2) const {
    fn_t f1 = m_msc->m_cp_fns[p1.first], f2 = m_msc->m_cp_fns[p1.second];
    fn_t f3 = m_msc->m_cp_fns[p2.first], f4 = m_msc->m_cp_fns[p2.second];
    fn_t d1 = std::abs(f2 - f1), d2 = std::abs(f4 - f3);
    if (d1 != d2) return d1 > d2;
    cellid_t c1 = m_msc->m_cps[p1.first]->cellid, c2 = m_msc->m_cps[p1.second]->cellid;
    cellid_t c3 = m_msc->m_cps[p1.first]->cellid, c4 = m_msc->m_cps[p1.second]->cellid;
    d1 = (fn_t)((c1[0] - c2[0]) * (c1[0] - c2[0]) + (c1[1] - c2[1]) * (c1[1] - c2[1]));
    d2 = (fn_t)((c3[0] - c4[0]) * (c3[0] - c4[0]) + (c3[1] - c4[1]) * (c3[1] - c4[1]));
    if (d1 != d2) return d1 > d2;
    if (c1 > c2) std::swap(c1, c2);
    if (c3 > c4) std::swap(c3, c4);
    if (c1 != c3) return c1 > c3;
    return c2 > c4;
  }
};

// grid_mscomplex.cpp:384-512
void GridMSComplex::simplify(crit_idx_pair_list_t &canc_pairs_list, double thresh) {
  persistence_comparator_t comp(this);
  std::priority_queue<crit_idx_pair_t, crit_idx_pair_list_t, persistence_comparator_t> pq(comp);

  fn_t max_val = std::numeric_limits<fn_t>::min();  // FLT_MIN, as at :399
  if (g_plant == 4) max_val = std::numeric_limits<fn_t>::lowest();
  fn_t min_val = std::numeric_limits<fn_t>::max();

  for (uint i = 0; i < m_cps.size(); ++i) {
    critpt_t *cp = m_cps[i];
    if (GridDataset::getCellDim(cp->cellid) == 1) {
      critpt_t::conn_t *acdc[] = {&cp->asc, &cp->des};
      for (uint j = 0; j < 2; ++j) {
        if (acdc[j]->size() != 2) continue;
        if (*(acdc[j]->begin()) == *(++acdc[j]->begin())) cp->isOnStrangulationPath = true;
      }
    }
    max_val = std::max(max_val, m_cp_fns[i]);
    min_val = std::min(min_val, m_cp_fns[i]);
    for (auto it = cp->des.begin(); it != cp->des.end(); ++it) pq.push(std::make_pair(i, *it));
  }

  crit_idx_pair_list_t resubmit;
  fn_t max_persistence = max_val - min_val;

  while (pq.size() != 0) {
    crit_idx_pair_t cpair = pq.top();
    pq.pop();
    uint v1 = cpair.first, v2 = cpair.second;
    critpt_t *cp1 = m_cps[v1], *cp2 = m_cps[v2];
    fn_t pers = std::abs(m_cp_fns[v1] - m_cp_fns[v2]);
    if ((double)pers / (double)max_persistence > thresh) break;
    if (GridDataset::getCellDim(cp2->cellid) == 1) {
      std::swap(cp1, cp2);
      std::swap(v1, v2);
    }
    uint cp2_dim = GridDataset::getCellDim(cp2->cellid);
    if (cp1->isCancelled || cp2->isCancelled) continue;
    if (m_rect.isOnBoundry(cp1->cellid) && !m_rect.isOnBoundry(cp2->cellid)) continue;
    if (!m_rect.isOnBoundry(cp1->cellid) && m_rect.isOnBoundry(cp2->cellid)) continue;
    critpt_t::conn_t *cp2_acdc[] = {&cp2->asc, &cp2->des};
    if (cp1->isOnStrangulationPath && cp2_acdc[cp2_dim / 2]->size() != 1) {
      resubmit.push_back(cpair);
      continue;
    }
    if (cp1->isOnStrangulationPath && m_rect.isOnBoundry(cp1->cellid)) continue;

    std::vector<uint> new_edges;
    cancelPairs(this, v1, v2, &new_edges);
    cp1->isBoundryCancelable = true;
    cp2->isBoundryCancelable = true;
    cp1->pair_idx = v2;
    cp2->pair_idx = v1;
    canc_pairs_list.push_back(cpair);
    for (size_t i = 0; i < new_edges.size(); i += 2) pq.push(std::make_pair(new_edges[i], new_edges[i + 1]));
    for (size_t i = 0; i < resubmit.size(); i++) pq.push(resubmit[i]);
    resubmit.clear();
  }
}

// grid_mscomplex.cpp:514-526
void GridMSComplex::un_simplify(const crit_idx_pair_list_t &lst) {
  for (auto it = lst.rbegin(); it != lst.rend(); ++it) uncancel_pairs(this, it->first, it->second);
}

// grid_mscomplex.cpp:528-535
void GridMSComplex::simplify_un_simplify(double thresh) {
  crit_idx_pair_list_t lst;
  simplify(lst, thresh);
  un_simplify(lst);
}

// ---------------------------------------------------------------- DataManager

DataManager::DataManager(const fn_t *data, int sx, int sy, int levels, int lx, int ly, double thr, bool keep)
    : m_data(data), m_size_x(sx), m_size_y(sy), m_simp_tresh(thr), m_keep(keep) {
  m_two_d = (lx > 0 || ly > 0);
  m_lx = lx;
  m_ly = ly;
  m_num_levels = m_two_d ? lx + ly : levels;
  // grid_datamanager.cpp:40-48
  rect_t r(cellid_t(0, 0), cellid_t(2 * (sx - 1), 2 * (sy - 1)));
  rect_t e = r;
  createPieces(r, e, m_num_levels, lx, ly);
  // A tiling finer than the grid gives pieces whose ext rect (the split
  // line plus one vertex of halo) leaves the grid; readDataAndInit would
  // copy past the end of the buffer for them (undefined behaviour in the
  // reference), so the restatement refuses such tilings up front.
  for (Piece *p : m_pieces) {
    const rect_t &pe = p->dataset->m_ext_rect;
    if (pe.bl[0] < 0 || pe.bl[1] < 0 || pe.tr[0] > r.tr[0] || pe.tr[1] > r.tr[1]) {
      // the destructor does not run for a throwing constructor
      for (Piece *q : m_pieces) {
        delete q->dataset;
        delete q->msgraph;
        delete q;
      }
      m_pieces.clear();
      throw std::invalid_argument("tiling is finer than the grid: a piece's ext rect leaves the grid");
    }
  }
}

DataManager::~DataManager() {
  for (Piece *p : m_pieces) {
    delete p->dataset;
    delete p->msgraph;
    delete p;
  }
}

// grid_datamanager.cpp:53-89.  With a 2-D tiling the split axis is chosen
// per level (y while it has more levels left than x, else x); the split
// position and the 2-cell halo follow the reference rule on that axis.
void DataManager::createPieces(rect_t r, rect_t e, int level, int lx, int ly) {
  if (level == 0) {
    Piece *dp = new Piece;
    dp->dataset = new GridDataset(r, e);
    dp->msgraph = new GridMSComplex(r, e);
    m_pieces.push_back(dp);
    return;
  }
  int dim = 1;
  if (m_two_d) dim = (ly >= lx && ly > 0) ? 1 : 0;
  cellid_t s = r.size();
  cellid_t tr1 = r.tr, bl1 = r.bl;
  tr1[dim] -= 2 * (s[dim] / 4);
  bl1[dim] = tr1[dim];
  rect_t r1(r.bl, tr1), r2(bl1, r.tr);
  cellid_t tr2 = e.tr, bl2 = e.bl;
  tr2[dim] = tr1[dim] + 2;
  bl2[dim] = bl1[dim] - 2;
  rect_t e1(e.bl, tr2), e2(bl2, e.tr);
  int nlx = lx, nly = ly;
  if (m_two_d) {
    if (dim == 1)
      nly--;
    else
      nlx--;
  }
  createPieces(r1, e1, level - 1, nlx, nly);
  createPieces(r2, e2, level - 1, nlx, nly);
}

// readDataAndInit (grid_datamanager.cpp:530-577): the piece reads its ext
// window of the raw x-fastest grid.
void DataManager::initPiece(Piece *dp) {
  const rect_t &e = dp->dataset->m_ext_rect;
  int x0 = e.bl[0] / 2, y0 = e.bl[1] / 2, x1 = e.tr[0] / 2, y1 = e.tr[1] / 2;
  int w = x1 - x0 + 1, h = y1 - y0 + 1;
  dp->data.resize((size_t)w * h);
  for (int j = 0; j < h; ++j)
    memcpy(&dp->data[(size_t)j * w], m_data + (size_t)(y0 + j) * m_size_x + x0, sizeof(fn_t) * w);
  dp->dataset->init(dp->data.data());
}

void DataManager::snapshot(std::vector<std::vector<uint32_t>> &dst) {
  if (!m_keep) return;
  dst.resize(m_pieces.size());
  for (size_t i = 0; i < m_pieces.size(); ++i) {
    if (m_pieces[i]->msgraph)
      m_pieces[i]->msgraph->dump(dst[i]);
    else
      dst[i].clear();
  }
}

// f(0..n-1) on up to nthreads threads (items taken in index order)
static void par_for(uint n, int nthreads, const std::function<void(uint)> &f) {
  if (nthreads <= 1 || n <= 1) {
    for (uint i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<uint> next(0);
  std::vector<std::exception_ptr> errs(n);
  auto worker = [&]() {
    for (uint i; (i = next.fetch_add(1)) < n;) {
      try {
        f(i);
      } catch (...) {
        errs[i] = std::current_exception();
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads && (uint)t < n; ++t) th.emplace_back(worker);
  for (auto &t : th) t.join();
  for (auto &e : errs)
    if (e) std::rethrow_exception(e);
}

// GridDataManager::GridDataManager compute order (grid_datamanager.cpp:
// 643-660): computeSubdomainMsgraphs, mergePiecesUp, root
// simplify_un_simplify, mergePiecesDown, collectSubdomainManifolds.  With
// nthreads == 1 this is the -st order; otherwise the leaf phases and each
// merge-tree level run their independent pieces in parallel
// (grid_datamanager.cpp:213-238, 258-306, 308-357).
void DataManager::run(PipelineSnapshots *snaps, int nthreads) {
  PipelineSnapshots local;
  PipelineSnapshots &S = snaps ? *snaps : local;
  uint num_leafs = 1u << m_num_levels;

  double t0 = now_s();
  // computeSubdomainMsgraphs / computeMsGraph (:121-140, :493-528)
  par_for(num_leafs, nthreads, [&](uint i) {
    Piece *dp = m_pieces[i];
    initPiece(dp);
    dp->dataset->work();
    dp->dataset->writeout_connectivity(dp->msgraph);
  });
  double t1 = now_s();
  S.t_compute = t1 - t0;
  snapshot(S.after_compute);

  // mergePiecesUp (:258-306), single-threaded order
  for (uint i = 0; i + 1 < num_leafs; ++i) m_pieces.push_back(new Piece);
  // pairs (j, j+1) of one level are merged into num_leafs + j/2; a level's
  // pairs start where the previous level's outputs start
  for (uint lo = 0, w = num_leafs; w > 1; lo += w, w /= 2) {
    par_for(w / 2, nthreads, [&](uint k) {
      uint j = lo + 2 * k;
      Piece *dp1 = m_pieces[j], *dp2 = m_pieces[j + 1], *dp = m_pieces[num_leafs + j / 2];
      if (dp1->level != dp2->level) throw std::logic_error("dps must have same level");
      dp->level = dp1->level + 1;
      dp->msgraph = GridMSComplex::merge_up(*dp1->msgraph, *dp2->msgraph);
    });
  }
  double t2 = now_s();
  S.t_merge_up = t2 - t1;
  snapshot(S.after_merge_up);
  if (g_stop_after < 2) return;

  // root simplification (:651-652)
  if (m_simp_tresh > 0.0) m_pieces.back()->msgraph->simplify_un_simplify(m_simp_tresh);
  double t3 = now_s();
  S.t_simplify = t3 - t2;
  snapshot(S.after_simplify);
  if (g_stop_after < 3) return;

  // mergePiecesDown (:308-357), single-threaded order.  With zero levels the
  // reference's loop bound shifts by -1; no iteration runs.
  uint num_graphs = (1u << (m_num_levels + 1)) - 1;
  if (m_num_levels >= 1) {
    // node j's children are 2j and 2j+1: one level [a, 2a) at a time
    const uint jn = 1u << (m_num_levels - 1);
    for (uint a = 1; a < jn; a *= 2) {
      const uint b = std::min(2 * a, jn);
      par_for(b - a, nthreads, [&](uint k) {
        uint j = a + k;
        Piece *dp = m_pieces[num_graphs - j];
        Piece *dp1 = m_pieces[num_graphs - 2 * j];
        Piece *dp2 = m_pieces[num_graphs - 2 * j - 1];
        dp->msgraph->merge_down(*dp1->msgraph, *dp2->msgraph);
      });
    }
  }
  // finalMergeDownPiecesInRange (:359-395), run for even leaves in
  // collectSubdomainManifolds (:466-485)
  if (m_num_levels != 0) {
    par_for(num_leafs / 2, nthreads, [&](uint k) {
      uint j = 2 * k;
      Piece *dp = m_pieces[num_leafs + j / 2];
      dp->msgraph->merge_down(*m_pieces[j]->msgraph, *m_pieces[j + 1]->msgraph);
    });
  }
  double t4 = now_s();
  S.t_merge_down = t4 - t3;
  snapshot(S.after_merge_down);
  if (g_stop_after < 4) return;

  // collectManifold (:397-422): re-init (flags reset, critical list kept)
  // and re-run work(), then postMergeFillDiscs.
  par_for(num_leafs, nthreads, [&](uint i) {
    Piece *dp = m_pieces[i];
    initPiece(dp);
    dp->dataset->work();
    dp->dataset->postMergeFillDiscs(dp->msgraph);
  });
  double t5 = now_s();
  S.t_manifolds = t5 - t4;
  snapshot(S.after_discs);
}

}  // namespace orc
