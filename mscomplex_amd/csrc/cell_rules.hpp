// Local discrete-gradient rules for one tile of the doubled cell grid.
//
// The reference assigns the gradient with a raster-order greedy loop
// (dmsc/src/grid_dataset.cpp:1743-1820 over lowestPairableCoFacet,
// :1314-1373).  That loop's result does not depend on the raster order: a
// cell c is paired with its lowest usable cofacet cand(c) when one exists,
// and otherwise with the facet f whose cand(f) == c, if any.  (cand(v) = e
// and cand(e) = q cannot both hold: e would have to contain the maximum
// vertex M of q and the larger of M's two neighbours in q, while cand(v)
// picks v's lowest lower neighbour, which is the smaller one.)  So every
// cell's pairing is a function of a 4x4 vertex window, which is what these
// functions compute.  They are __host__ __device__ so the same code runs in
// the HIP kernels and in CPU tests against the oracle.
//
// Coordinates here are LOCAL to the tile's ext rect: vertex (i, j) is cell
// (2i, 2j); ext spans cells [0, cw) x [0, ch) with cw = 2 vx - 1.  The
// interior rect (the reference's m_rect) is [rx0, rx1] x [ry0, ry1].
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define MSCX_HD __host__ __device__ __forceinline__
#else
#define MSCX_HD inline
#endif

namespace mscx {

// Gradient code, one byte per cell:
//   bits 0-1: the reference's cell flags (1 = paired, 2 = critical; 3 =
//             boundary-cancellable pair, grid_dataset.h:40-45)
//   bits 2-4: direction of the paired cell (0 none, 1 +x, 2 -x, 3 +y, 4 -y)
enum : uint8_t { CF_PAIRED = 1, CF_CRITICAL = 2 };
enum : int { DIR_NONE = 0, DIR_PX = 1, DIR_NX = 2, DIR_PY = 3, DIR_NY = 4 };

MSCX_HD int dir_dx(int d) { return d == DIR_PX ? 1 : (d == DIR_NX ? -1 : 0); }
MSCX_HD int dir_dy(int d) { return d == DIR_PY ? 1 : (d == DIR_NY ? -1 : 0); }
MSCX_HD int dir_flip(int d) { return d == 0 ? 0 : (d == DIR_PX ? DIR_NX : d == DIR_NX ? DIR_PX : d == DIR_PY ? DIR_NY : DIR_PY); }
MSCX_HD uint8_t code_flags(uint8_t c) { return c & 3; }
MSCX_HD int code_dir(uint8_t c) { return (c >> 2) & 7; }

struct TileGeom {
  int vx, vy;              // vertex extents of the ext window
  int cw, ch;              // cell extents (2 vx - 1, 2 vy - 1)
  int rx0, ry0, rx1, ry1;  // interior rect, local cell coords (inclusive)
};

// Total order on vertices: value, then cell coordinates x-major
// (grid_dataset.h:158-167).  -0.0 and +0.0 compare equal and fall through
// to the coordinate tie-break.  Inputs are NaN-free (checked at the API).
MSCX_HD uint32_t fkey(float f) {
  union {
    float f;
    uint32_t u;
  } v;
  v.f = f;
  uint32_t u = (f == 0.0f) ? 0u : v.u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
MSCX_HD uint64_t vkey(uint32_t fk, int i, int j) {
  return ((uint64_t)fk << 28) | ((uint64_t)(uint32_t)i << 14) | (uint64_t)(uint32_t)j;
}

MSCX_HD bool in_ext_v(const TileGeom &g, int i, int j) { return i >= 0 && j >= 0 && i < g.vx && j < g.vy; }
MSCX_HD bool in_rect(const TileGeom &g, int cx, int cy) {
  return cx >= g.rx0 && cx <= g.rx1 && cy >= g.ry0 && cy <= g.ry1;
}
// on the ext rect boundary (the reference's "true boundary", grid_dataset.cpp:1698-1701)
MSCX_HD bool on_ext_bndry(const TileGeom &g, int cx, int cy) {
  return cx == 0 || cy == 0 || cx == g.cw - 1 || cy == g.ch - 1;
}

// Vertex key with the reference's order (value, then x, then y) kept as
// separate fields: when the kernels compare neighbours whose relative
// positions are fixed, the position tie-break folds to a constant and the
// comparison becomes one 32-bit compare.  Same order as vkey().
struct VKey {
  uint32_t f;
  int i, j;
};
MSCX_HD bool operator<(const VKey &a, const VKey &b) {
  return a.f != b.f ? a.f < b.f : (a.i != b.i ? a.i < b.i : a.j < b.j);
}
MSCX_HD bool operator>(const VKey &a, const VKey &b) { return b < a; }
MSCX_HD bool operator==(const VKey &a, const VKey &b) { return a.f == b.f && a.i == b.i && a.j == b.j; }
MSCX_HD bool operator!=(const VKey &a, const VKey &b) { return !(a == b); }

template <class T>
MSCX_HD void cswap(T &x, T &y) {
  if (y < x) {
    T t = x;
    x = y;
    y = t;
  }
}
template <class T>
MSCX_HD void sort4(T &a, T &b, T &c, T &d) {
  cswap(a, b);
  cswap(c, d);
  cswap(a, c);
  cswap(b, d);
  cswap(b, c);
}

// Lexicographic "quad p < quad q" on ascending-sorted vertex keys
// (grid_dataset.cpp:1519-1536 for dim-2 cells).
template <class T>
MSCX_HD bool quad_lt(T p0, T p1, T p2, T p3, T q0, T q1, T q2, T q3) {
  sort4(p0, p1, p2, p3);
  sort4(q0, q1, q2, q3);
  if (p0 != q0) return p0 < q0;
  if (p1 != q1) return p1 < q1;
  if (p2 != q2) return p2 < q2;
  return p3 < q3;
}

// cand(v): the lowest usable edge of vertex (i, j) (lowestPairableCoFacet
// for a dim-0 cell).  An edge is usable when v is its larger vertex and,
// for v on the ext boundary, the edge also lies on it.  Returns DIR_*.
template <class KF>
MSCX_HD int vertex_cand(const KF &K, const TileGeom &g, int i, int j) {
  const auto kv = K(i, j);
  bool tb = (i == 0 || j == 0 || i == g.vx - 1 || j == g.vy - 1);
  int best = DIR_NONE;
  auto bk = kv;  // meaningful once best != DIR_NONE
  // +x / -x: horizontal edge on the boundary iff its row is a boundary row
  bool hrow_tb = (j == 0 || j == g.vy - 1);
  bool vcol_tb = (i == 0 || i == g.vx - 1);
  if (i + 1 < g.vx && (!tb || hrow_tb)) {
    const auto k = K(i + 1, j);
    if (k < kv && (best == DIR_NONE || k < bk)) { bk = k; best = DIR_PX; }
  }
  if (i - 1 >= 0 && (!tb || hrow_tb)) {
    const auto k = K(i - 1, j);
    if (k < kv && (best == DIR_NONE || k < bk)) { bk = k; best = DIR_NX; }
  }
  if (j + 1 < g.vy && (!tb || vcol_tb)) {
    const auto k = K(i, j + 1);
    if (k < kv && (best == DIR_NONE || k < bk)) { bk = k; best = DIR_PY; }
  }
  if (j - 1 >= 0 && (!tb || vcol_tb)) {
    const auto k = K(i, j - 1);
    if (k < kv && (best == DIR_NONE || k < bk)) { bk = k; best = DIR_NY; }
  }
  return best;
}

// cand(e) for the horizontal edge (2i+1, 2j) between vertices (i,j) and
// (i+1,j).  Its cofacets are the quads above (+y) and below (-y).  A quad
// is usable when e is its largest edge, i.e. e joins the quad's maximum
// vertex to the larger of that vertex's two quad neighbours.  Edges on the
// ext boundary have no usable quad (quads are never on it).
template <class KF>
MSCX_HD int hedge_cand(const KF &K, const TileGeom &g, int i, int j) {
  if (j == 0 || j == g.vy - 1) return DIR_NONE;
  const auto a = K(i, j), b = K(i + 1, j);
  bool a_hi = a > b;
  const auto hi = a_hi ? a : b, lo = a_hi ? b : a;
  int hi_i = a_hi ? i : i + 1, lo_i = a_hi ? i + 1 : i;
  const auto u_hi = K(hi_i, j + 1), u_lo = K(lo_i, j + 1);
  const auto d_hi = K(hi_i, j - 1), d_lo = K(lo_i, j - 1);
  bool up = (u_hi < lo) && (u_lo < hi);
  bool dn = (d_hi < lo) && (d_lo < hi);
  if (up && dn) return quad_lt(lo, hi, u_hi, u_lo, lo, hi, d_hi, d_lo) ? DIR_PY : DIR_NY;
  return up ? DIR_PY : (dn ? DIR_NY : DIR_NONE);
}

// cand(e) for the vertical edge (2i, 2j+1) between (i,j) and (i,j+1); its
// cofacets are the quads right (+x) and left (-x).
template <class KF>
MSCX_HD int vedge_cand(const KF &K, const TileGeom &g, int i, int j) {
  if (i == 0 || i == g.vx - 1) return DIR_NONE;
  const auto a = K(i, j), b = K(i, j + 1);
  bool a_hi = a > b;
  const auto hi = a_hi ? a : b, lo = a_hi ? b : a;
  int hi_j = a_hi ? j : j + 1, lo_j = a_hi ? j + 1 : j;
  const auto r_hi = K(i + 1, hi_j), r_lo = K(i + 1, lo_j);
  const auto l_hi = K(i - 1, hi_j), l_lo = K(i - 1, lo_j);
  bool rt = (r_hi < lo) && (r_lo < hi);
  bool lf = (l_hi < lo) && (l_lo < hi);
  if (rt && lf) return quad_lt(lo, hi, r_hi, r_lo, lo, hi, l_hi, l_lo) ? DIR_PX : DIR_NX;
  return rt ? DIR_PX : (lf ? DIR_NX : DIR_NONE);
}

// A paired cell x lies on a side of the interior rect that is not a side of
// the ext rect; such a cell has an exterior cofacet, so the reference marks
// it and its pair critical (grid_dataset.cpp:1767-1819).
MSCX_HD bool bndry_cancellable_site(const TileGeom &g, int cx, int cy) {
  if (!in_rect(g, cx, cy)) return false;
  return (cy == g.ry0 && g.ry0 > 0) || (cy == g.ry1 && g.ry1 < g.ch - 1) || (cx == g.rx0 && g.rx0 > 0) ||
         (cx == g.rx1 && g.rx1 < g.cw - 1);
}

// Gradient code of local cell (cx, cy); K(i, j) returns the key of vertex
// (i, j) for every vertex of the 4x4 window around the cell that lies in
// the ext rect.
template <class KF>
MSCX_HD uint8_t cell_code(const KF &K, const TileGeom &g, int cx, int cy) {
  int dx = cx & 1, dy = cy & 1;
  bool inr = in_rect(g, cx, cy);
  int dir = DIR_NONE;
  if (!dx && !dy) {  // vertex: pairs only upward
    if (inr) dir = vertex_cand(K, g, cx >> 1, cy >> 1);
  } else if (dx && !dy) {  // horizontal edge
    int i = cx >> 1, j = cy >> 1;
    if (inr) dir = hedge_cand(K, g, i, j);
    if (dir == DIR_NONE) {
      bool a_hi = K(i, j) > K(i + 1, j);
      int hi_i = a_hi ? i : i + 1;
      if (in_rect(g, 2 * hi_i, cy) && vertex_cand(K, g, hi_i, j) == (a_hi ? DIR_PX : DIR_NX))
        dir = a_hi ? DIR_NX : DIR_PX;
    }
  } else if (!dx && dy) {  // vertical edge
    int i = cx >> 1, j = cy >> 1;
    if (inr) dir = vedge_cand(K, g, i, j);
    if (dir == DIR_NONE) {
      bool a_hi = K(i, j) > K(i, j + 1);
      int hi_j = a_hi ? j : j + 1;
      if (in_rect(g, cx, 2 * hi_j) && vertex_cand(K, g, i, hi_j) == (a_hi ? DIR_PY : DIR_NY))
        dir = a_hi ? DIR_NY : DIR_PY;
    }
  } else {  // quad: paired with its largest edge when that edge picks it
    int i = cx >> 1, j = cy >> 1;
    const auto k00 = K(i, j), k10 = K(i + 1, j), k01 = K(i, j + 1), k11 = K(i + 1, j + 1);
    // maximum corner and its larger quad neighbour give the largest edge
    int edir;  // direction from the quad to that edge
    if (k00 > k10 && k00 > k01 && k00 > k11) edir = (k10 > k01) ? DIR_NY : DIR_NX;
    else if (k10 > k01 && k10 > k11) edir = (k00 > k11) ? DIR_NY : DIR_PX;
    else if (k01 > k11) edir = (k00 > k11) ? DIR_NX : DIR_PY;
    else edir = (k10 > k01) ? DIR_PX : DIR_PY;
    int ex = cx + dir_dx(edir), ey = cy + dir_dy(edir);
    if (in_rect(g, ex, ey)) {
      int ec = (ex & 1) ? hedge_cand(K, g, ex >> 1, ey >> 1) : vedge_cand(K, g, ex >> 1, ey >> 1);
      if (ec == dir_flip(edir)) dir = edir;
    }
  }
  uint8_t flags;
  if (dir != DIR_NONE) {
    flags = CF_PAIRED;
    int px = cx + dir_dx(dir), py = cy + dir_dy(dir);
    if (bndry_cancellable_site(g, cx, cy) || bndry_cancellable_site(g, px, py)) flags |= CF_CRITICAL;
  } else {
    flags = inr ? CF_CRITICAL : 0;
  }
  return (uint8_t)(flags | (dir << 2));
}

// The same code computed from candidates that are evaluated once per cell:
// VC(i, j) = vertex_cand of vertex (i, j), HC(i, j) = hedge_cand of edge
// (2i+1, 2j), EC(i, j) = vedge_cand of edge (2i, 2j+1), each DIR_NONE for a
// cell outside the interior rect (exterior cells never pick a pair).  The
// HIP kernel stores these in LDS; cell_code above is the per-cell form used
// to check this one.
template <class KF, class VCF, class HCF, class ECF>
MSCX_HD uint8_t code_from_cands(const KF &K, const VCF &VC, const HCF &HC, const ECF &EC, const TileGeom &g,
                                int cx, int cy) {
  const int dx = cx & 1, dy = cy & 1;
  const int i = cx >> 1, j = cy >> 1;
  int dir = DIR_NONE;
  if (!dx && !dy) {
    dir = VC(i, j);
  } else if (dx && !dy) {
    dir = HC(i, j);
    if (dir == DIR_NONE) {
      const bool a_hi = K(i, j) > K(i + 1, j);
      if (a_hi ? VC(i, j) == DIR_PX : VC(i + 1, j) == DIR_NX) dir = a_hi ? DIR_NX : DIR_PX;
    }
  } else if (!dx && dy) {
    dir = EC(i, j);
    if (dir == DIR_NONE) {
      const bool a_hi = K(i, j) > K(i, j + 1);
      if (a_hi ? VC(i, j) == DIR_PY : VC(i, j + 1) == DIR_NY) dir = a_hi ? DIR_NY : DIR_PY;
    }
  } else {
    const auto k00 = K(i, j), k10 = K(i + 1, j), k01 = K(i, j + 1), k11 = K(i + 1, j + 1);
    int edir;
    if (k00 > k10 && k00 > k01 && k00 > k11) edir = (k10 > k01) ? DIR_NY : DIR_NX;
    else if (k10 > k01 && k10 > k11) edir = (k00 > k11) ? DIR_NY : DIR_PX;
    else if (k01 > k11) edir = (k00 > k11) ? DIR_NX : DIR_PY;
    else edir = (k10 > k01) ? DIR_PX : DIR_PY;
    int ec;
    switch (edir) {
      case DIR_NY: ec = HC(i, j); break;
      case DIR_PY: ec = HC(i, j + 1); break;
      case DIR_NX: ec = EC(i, j); break;
      default: ec = EC(i + 1, j); break;
    }
    if (ec == dir_flip(edir)) dir = edir;
  }
  const bool inr = in_rect(g, cx, cy);
  uint8_t flags;
  if (dir != DIR_NONE) {
    flags = CF_PAIRED;
    const int px = cx + dir_dx(dir), py = cy + dir_dy(dir);
    if (bndry_cancellable_site(g, cx, cy) || bndry_cancellable_site(g, px, py)) flags |= CF_CRITICAL;
  } else {
    flags = inr ? CF_CRITICAL : 0;
  }
  return (uint8_t)(flags | (dir << 2));
}

}  // namespace mscx
