// tests/pins.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Checks of a Morse-Smale complex dump (the format of mscx_complex_dump /
// the oracle's dump) that the mathematics fixes, independent of how the
// complex was computed.  Used on the oracle's output (CPU tests) and on the
// HIP pipeline's output (GPU tests, at sizes the oracle cannot reach).
//
//   pin_nodes      per-node arrays (cell, fn bits, pair, flags) of a dump
//   pin_canon      canonical form of the live nodes: sorted by cell id, arcs
//                  as sorted multisets of cell ids (index-free, so complexes
//                  of different tilings compare)
//   pin_homology   Morse homology over Z/2 of the complex restricted to a
//                  node mask: ranks of the boundary maps d1 (saddle -> min)
//                  and d2 (max -> saddle), d1 d2 = 0 (mod 2), and the arc
//                  weights of every saddle (two descending paths; two
//                  ascending, one on the domain boundary)
//   pin_discs      the ascending discs of the surviving minima partition the
//                  rect's vertices and the descending discs of the surviving
//                  maxima partition its quads
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

struct Node {
  uint32_t cell, fn, pair, flags;
  const uint32_t *asc, *des, *ad, *dd;
  uint32_t na, nd, nad, ndd;
};

struct Dump {
  int32_t rect[4], ext[4];
  std::vector<Node> nodes;
  bool ok = false;
};

bool parse(const uint32_t *d, size_t n, Dump &out) {
  if (n < 9) return false;
  const uint32_t k = d[0];
  for (int i = 0; i < 4; ++i) out.rect[i] = (int32_t)d[1 + i], out.ext[i] = (int32_t)d[5 + i];
  out.nodes.resize(k);
  size_t p = 9;
  for (uint32_t i = 0; i < k; ++i) {
    if (p + 8 > n) return false;
    Node &c = out.nodes[i];
    c.cell = d[p], c.fn = d[p + 1], c.pair = d[p + 2], c.flags = d[p + 3];
    c.na = d[p + 4], c.nd = d[p + 5], c.nad = d[p + 6], c.ndd = d[p + 7];
    p += 8;
    size_t need = (size_t)c.na + c.nd + c.nad + c.ndd;
    if (p + need > n) return false;
    c.asc = d + p;
    p += c.na;
    c.des = d + p;
    p += c.nd;
    c.ad = d + p;
    p += c.nad;
    c.dd = d + p;
    p += c.ndd;
  }
  out.ok = (p == n);
  return out.ok;
}

inline int cx(uint32_t c) { return (int16_t)(c & 0xFFFF); }
inline int cy(uint32_t c) { return (int16_t)(c >> 16); }
inline int dim(uint32_t c) { return (c & 1) + ((c >> 16) & 1); }

struct DSU {
  std::vector<uint32_t> p;
  explicit DSU(size_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
  uint32_t find(uint32_t x) {
    while (p[x] != x) {
      p[x] = p[p[x]];
      x = p[x];
    }
    return x;
  }
  bool unite(uint32_t a, uint32_t b) {
    a = find(a), b = find(b);
    if (a == b) return false;
    p[a] = b;
    return true;
  }
};

}  // namespace

extern "C" {

// Node count of a dump (-1 if malformed); per-node arrays are written when
// the pointers are not NULL.
long pin_nodes(const uint32_t *d, size_t n, uint32_t *cell, uint32_t *fn, uint32_t *pair, uint32_t *flags,
               uint32_t *counts4) {
  Dump D;
  if (!parse(d, n, D)) return -1;
  for (size_t i = 0; i < D.nodes.size(); ++i) {
    const Node &c = D.nodes[i];
    if (cell) cell[i] = c.cell;
    if (fn) fn[i] = c.fn;
    if (pair) pair[i] = c.pair;
    if (flags) flags[i] = c.flags;
    if (counts4) {
      counts4[4 * i] = c.na, counts4[4 * i + 1] = c.nd, counts4[4 * i + 2] = c.nad, counts4[4 * i + 3] = c.ndd;
    }
  }
  return (long)D.nodes.size();
}

// Canonical form of the nodes selected by mask (mask == NULL: the nodes
// without the cancelled flag): per node in cell-id order, [cell, fn bits,
// n_asc, n_des, asc cells (sorted), des cells (sorted)].  Returns the word
// count (the words are written when cap suffices), -1 for a malformed dump,
// -2 when an arc of a selected node leads to an unselected or missing node.
long pin_canon(const uint32_t *d, size_t n, const uint8_t *mask, uint32_t *out, size_t cap) {
  Dump D;
  if (!parse(d, n, D)) return -1;
  const size_t K = D.nodes.size();
  std::vector<uint8_t> sel(K);
  for (size_t i = 0; i < K; ++i) sel[i] = mask ? mask[i] != 0 : !(D.nodes[i].flags & 1u);
  std::vector<uint32_t> order;
  for (size_t i = 0; i < K; ++i)
    if (sel[i]) order.push_back((uint32_t)i);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return D.nodes[a].cell < D.nodes[b].cell; });
  std::vector<uint32_t> w, tmp;
  for (uint32_t i : order) {
    const Node &c = D.nodes[i];
    w.push_back(c.cell);
    w.push_back(c.fn);
    w.push_back(c.na);
    w.push_back(c.nd);
    for (int s = 0; s < 2; ++s) {
      const uint32_t *a = s ? c.des : c.asc;
      uint32_t m = s ? c.nd : c.na;
      tmp.clear();
      for (uint32_t j = 0; j < m; ++j) {
        if (a[j] >= K || !sel[a[j]]) return -2;
        tmp.push_back(D.nodes[a[j]].cell);
      }
      std::sort(tmp.begin(), tmp.end());
      w.insert(w.end(), tmp.begin(), tmp.end());
    }
  }
  if (out && cap >= w.size()) memcpy(out, w.data(), w.size() * 4);
  return (long)w.size();
}

// Morse homology over Z/2 of the nodes selected by mask, on the domain
// rect (the complex's rect, which must be the whole grid).  out[0..10]:
//   #min, #saddle, #max, rank d1, rank d2,
//   (max, min) pairs with an odd number of max -> saddle -> min paths,
//   arcs from a selected node to an unselected one,
//   saddles whose descending weight is not 2,
//   saddles whose ascending weight is not 2 (interior) or 1 (on the rect
//   boundary),
//   arcs that do not join adjacent dimensions,
//   asc/des arcs without the mirror arc (multiplicities differ).
// For a disc: rank d1 = #min - 1 and rank d2 = #max (H0 = Z/2, H1 = H2 = 0).
// Returns 0, or -1 for a malformed dump.
int pin_homology(const uint32_t *d, size_t n, const uint8_t *mask, int64_t *out) {
  Dump D;
  if (!parse(d, n, D)) return -1;
  const size_t K = D.nodes.size();
  const int32_t *R = D.rect;
  std::vector<int64_t> o(11, 0);
  std::vector<uint32_t> lid(K, UINT32_MAX);
  uint32_t nmin = 0, nmax = 0;
  for (size_t i = 0; i < K; ++i) {
    if (!mask[i]) continue;
    int dm = dim(D.nodes[i].cell);
    if (dm == 0) lid[i] = nmin++;
    if (dm == 2) lid[i] = nmax++;
    o[dm]++;
  }
  DSU mins(nmin), maxs(nmax + 1);  // maxs: index nmax is the outside
  int64_t r1 = 0, r2 = 0;
  std::vector<uint32_t> tmp;
  auto count_in = [](const uint32_t *a, uint32_t m, uint32_t v) {
    uint32_t k = 0;
    for (uint32_t j = 0; j < m; ++j) k += a[j] == v;
    return k;
  };
  for (size_t i = 0; i < K; ++i) {
    if (!mask[i]) continue;
    const Node &c = D.nodes[i];
    const int dm = dim(c.cell);
    for (int s = 0; s < 2; ++s) {
      const uint32_t *a = s ? c.des : c.asc;
      uint32_t m = s ? c.nd : c.na;
      for (uint32_t j = 0; j < m; ++j) {
        if (a[j] >= K || !mask[a[j]]) {
          o[6]++;
          continue;
        }
        const Node &t = D.nodes[a[j]];
        if (dim(t.cell) != dm + (s ? -1 : 1)) o[9]++;
        const uint32_t *b = s ? t.asc : t.des;
        uint32_t mb = s ? t.na : t.nd;
        if (count_in(b, mb, (uint32_t)i) != count_in(a, m, a[j])) o[10]++;
      }
    }
    if (dm != 1) continue;
    // descending: two paths to minima (counted with multiplicity)
    uint32_t wd = 0, wa = 0;
    tmp.clear();
    for (uint32_t j = 0; j < c.nd; ++j)
      if (c.des[j] < K && mask[c.des[j]]) tmp.push_back(lid[c.des[j]]), ++wd;
    if (wd != 2) o[7]++;
    if (tmp.size() == 2 && tmp[0] != tmp[1] && mins.unite(tmp[0], tmp[1])) ++r1;
    tmp.clear();
    for (uint32_t j = 0; j < c.na; ++j)
      if (c.asc[j] < K && mask[c.asc[j]]) tmp.push_back(lid[c.asc[j]]), ++wa;
    const int x = cx(c.cell), y = cy(c.cell);
    const bool onb = x == R[0] || x == R[2] || y == R[1] || y == R[3];
    if (wa != (onb ? 1u : 2u)) o[8]++;
    if (tmp.size() == 1) {
      if (maxs.unite(tmp[0], nmax)) ++r2;
    } else if (tmp.size() == 2 && tmp[0] != tmp[1]) {
      if (maxs.unite(tmp[0], tmp[1])) ++r2;
    }
  }
  o[3] = r1;
  o[4] = r2;
  // d1 d2 over Z/2: for each maximum, the parity of paths to each minimum
  std::vector<uint32_t> par(nmin, 0), touched;
  for (size_t i = 0; i < K; ++i) {
    if (!mask[i] || dim(D.nodes[i].cell) != 2) continue;
    const Node &c = D.nodes[i];
    touched.clear();
    for (uint32_t j = 0; j < c.nd; ++j) {
      uint32_t s = c.des[j];
      if (s >= K || !mask[s]) continue;
      const Node &t = D.nodes[s];
      for (uint32_t q = 0; q < t.nd; ++q) {
        uint32_t mn = t.des[q];
        if (mn >= K || !mask[mn] || lid[mn] == UINT32_MAX || dim(D.nodes[mn].cell) != 0) continue;
        par[lid[mn]] ^= 1u;
        touched.push_back(lid[mn]);
      }
    }
    for (uint32_t v : touched) {
      if (par[v]) o[5]++;
      par[v] = 0;
    }
  }
  for (int i = 0; i < 11; ++i) out[i] = o[i];
  return 0;
}

// Disc partition of the rect {x0,y0,x1,y1}: survivors are the nodes without
// the boundary-cancellable flag (postMergeFillDiscs' surviving extrema).
// out[0..5]: vertices of the rect not in any surviving minimum's ascending
// disc, vertices in more than one (or twice), disc cells that are not
// vertices of the rect; the same three for quads and the surviving maxima's
// descending discs.  out[6]: discs of non-surviving extrema that are not
// empty.  out[7]: discs not in raster order.  Returns 0, or -1 for a
// malformed dump.
int pin_discs(const uint32_t *d, size_t n, const int32_t *rect, int64_t *out) {
  Dump D;
  if (!parse(d, n, D)) return -1;
  const int x0 = rect[0], y0 = rect[1], x1 = rect[2], y1 = rect[3];
  const int64_t W = x1 - x0 + 1, H = y1 - y0 + 1;
  std::vector<uint8_t> seen((size_t)(W * H), 0);
  int64_t o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (const Node &c : D.nodes) {
    const int dm = dim(c.cell);
    if (dm == 1) continue;
    const uint32_t *a = dm == 0 ? c.ad : c.dd;
    const uint32_t m = dm == 0 ? c.nad : c.ndd;
    if (c.flags & 4u) {
      if (m) o[6]++;
      continue;
    }
    for (uint32_t j = 0; j < m; ++j) {
      const uint32_t v = a[j];
      const int x = cx(v), y = cy(v);
      if (j && a[j - 1] >= v) o[7]++;
      if (x < x0 || x > x1 || y < y0 || y > y1 || dim(v) != dm) {
        o[dm == 0 ? 2 : 5]++;
        continue;
      }
      uint8_t &s = seen[(size_t)(x - x0) + (size_t)(y - y0) * W];
      if (s) o[dm == 0 ? 1 : 4]++;
      s = 1;
    }
  }
  for (int y = y0; y <= y1; ++y)
    for (int x = x0; x <= x1; ++x) {
      const int dm = (x & 1) + (y & 1);
      if (dm == 1) continue;
      if (!seen[(size_t)(x - x0) + (size_t)(y - y0) * W]) o[dm == 0 ? 0 : 3]++;
    }
  for (int i = 0; i < 8; ++i) out[i] = o[i];
  return 0;
}

}  // extern "C"
