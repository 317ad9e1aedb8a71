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

// c ? a : b on keys.  For the split key it selects field by field: a
// select between two structs would otherwise be lowered through memory
// (a select of their addresses), which on the GPU means scratch.
template <class T>
MSCX_HD T ksel(bool c, const T &a, const T &b) {
  return c ? a : b;
}
MSCX_HD VKey ksel(bool c, const VKey &a, const VKey &b) {
  return VKey{c ? a.f : b.f, c ? a.i : b.i, c ? a.j : b.j};
}

template <class T>
MSCX_HD void cswap(T &x, T &y) {
  const bool s = y < x;
  const T lo = ksel(s, y, x), hi = ksel(s, x, y);
  x = lo;
  y = hi;
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
// The minimum is taken as a tournament over the neighbours in position
// order NX < NY < PY < PX, so a kernel with split keys compares each pair
// with the position tie-break already known.  INTERIOR: the caller knows
// (i, j) is not on the ext boundary, which removes every boundary test.
template <bool INTERIOR = false, class KF>
MSCX_HD int vertex_cand(const KF &K, const TileGeom &g, int i, int j) {
  const auto kv = K(i, j);
  const bool tb = !INTERIOR && (i == 0 || j == 0 || i == g.vx - 1 || j == g.vy - 1);
  // +x / -x: horizontal edge on the boundary iff its row is a boundary row
  const bool hx = !tb || (j == 0 || j == g.vy - 1);
  const bool vy = !tb || (i == 0 || i == g.vx - 1);
  const bool inx = INTERIOR || i - 1 >= 0, iny = INTERIOR || j - 1 >= 0;
  const bool ipy = INTERIOR || j + 1 < g.vy, ipx = INTERIOR || i + 1 < g.vx;
  // K is called at fixed offsets only (a neighbour outside the ext may be
  // read but its key is discarded)
  const auto knx = ksel(inx, K(i - 1, j), kv), kny = ksel(iny, K(i, j - 1), kv);
  const auto kpy = ksel(ipy, K(i, j + 1), kv), kpx = ksel(ipx, K(i + 1, j), kv);
  const bool unx = inx && hx && knx < kv, uny = iny && vy && kny < kv;
  const bool upy = ipy && vy && kpy < kv, upx = ipx && hx && kpx < kv;
  int m1 = unx ? DIR_NX : (uny ? DIR_NY : DIR_NONE);
  if (unx && uny && kny < knx) m1 = DIR_NY;
  int m2 = upy ? DIR_PY : (upx ? DIR_PX : DIR_NONE);
  if (upy && upx && kpx < kpy) m2 = DIR_PX;
  if (m1 == DIR_NONE) return m2;
  if (m2 == DIR_NONE) return m1;
  const auto k1 = ksel(m1 == DIR_NY, kny, knx);
  const auto k2 = ksel(m2 == DIR_PX, kpx, kpy);
  return (k2 < k1) ? m2 : m1;
}

// cand(e) for the horizontal edge (2i+1, 2j) between vertices (i,j) and
// (i+1,j).  Its cofacets are the quads above (+y) and below (-y).  A quad
// is usable when e is its largest edge, i.e. e joins the quad's maximum
// vertex to the larger of that vertex's two quad neighbours.  Edges on the
// ext boundary have no usable quad (quads are never on it).
template <bool INTERIOR = false, class KF>
MSCX_HD int hedge_cand(const KF &K, const TileGeom &g, int i, int j) {
  if (!INTERIOR && (j == 0 || j == g.vy - 1)) return DIR_NONE;
  const auto a = K(i, j), b = K(i + 1, j);
  bool a_hi = a > b;
  const auto hi = ksel(a_hi, a, b), lo = ksel(a_hi, b, a);
  const auto ua = K(i, j + 1), ub = K(i + 1, j + 1), da = K(i, j - 1), db = K(i + 1, j - 1);
  const auto u_hi = ksel(a_hi, ua, ub), u_lo = ksel(a_hi, ub, ua);
  const auto d_hi = ksel(a_hi, da, db), d_lo = ksel(a_hi, db, da);
  bool up = (u_hi < lo) && (u_lo < hi);
  bool dn = (d_hi < lo) && (d_lo < hi);
  if (up && dn) return quad_lt(lo, hi, u_hi, u_lo, lo, hi, d_hi, d_lo) ? DIR_PY : DIR_NY;
  return up ? DIR_PY : (dn ? DIR_NY : DIR_NONE);
}

// cand(e) for the vertical edge (2i, 2j+1) between (i,j) and (i,j+1); its
// cofacets are the quads right (+x) and left (-x).
template <bool INTERIOR = false, class KF>
MSCX_HD int vedge_cand(const KF &K, const TileGeom &g, int i, int j) {
  if (!INTERIOR && (i == 0 || i == g.vx - 1)) return DIR_NONE;
  const auto a = K(i, j), b = K(i, j + 1);
  bool a_hi = a > b;
  const auto hi = ksel(a_hi, a, b), lo = ksel(a_hi, b, a);
  const auto ra = K(i + 1, j), rb = K(i + 1, j + 1), la = K(i - 1, j), lb = K(i - 1, j + 1);
  const auto r_hi = ksel(a_hi, ra, rb), r_lo = ksel(a_hi, rb, ra);
  const auto l_hi = ksel(a_hi, la, lb), l_lo = ksel(a_hi, lb, la);
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
// HIP kernel keeps these in registers; cell_code above is the per-cell form
// used to check this one.  INTERIOR: the caller knows the cell and its
// neighbours lie strictly inside the interior rect (no exterior cells, no
// boundary-cancellable sites).
template <bool INTERIOR = false, class KF, class VCF, class HCF, class ECF>
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
    // candidates of the four edges, selected by value (a switch over the
    // accessor calls would select between the addresses of their sources)
    const int hc0 = HC(i, j), hc1 = HC(i, j + 1), ec0 = EC(i, j), ec1 = EC(i + 1, j);
    const int ec = edir == DIR_NY ? hc0 : edir == DIR_PY ? hc1 : edir == DIR_NX ? ec0 : ec1;
    if (ec == dir_flip(edir)) dir = edir;
  }
  const bool inr = INTERIOR || in_rect(g, cx, cy);
  uint8_t flags;
  if (dir != DIR_NONE) {
    flags = CF_PAIRED;
    const int px = cx + dir_dx(dir), py = cy + dir_dy(dir);
    if (!INTERIOR && (bndry_cancellable_site(g, cx, cy) || bndry_cancellable_site(g, px, py))) flags |= CF_CRITICAL;
  } else {
    flags = inr ? CF_CRITICAL : 0;
  }
  return (uint8_t)(flags | (dir << 2));
}

}  // namespace mscx

namespace mscx {

// ---------------------------------------------------------------------------
// Row-step form of the rules used by the gradient kernel.  A lane walks one
// vertex column i upward; for the current row j it holds the 4x4 window of
// value keys wRC = fkey of vertex (i - 1 + C, j - 1 + R) and the candidates
// of vertices (i, j), (i+1, j) and of the edge (i, j)-(i+1, j) carried from
// the previous row.  Every comparison below is between two vertices whose
// relative position is fixed, so the reference's position tie-break
// (grid_dataset.h:158-167: value, then x, then y) is a choice between <= and
// < on the 32-bit keys.  Each function equals the corresponding generic rule
// above (checked exhaustively on the CPU against cell_code).

constexpr uint32_t KEY_INF = 0xFFFFFFFFu;  // above every fkey() value (NaN-free input)

struct KeyWin {
  uint32_t w00, w01, w02, w03, w10, w11, w12, w13, w20, w21, w22, w23, w30, w31, w32, w33;
  int i, j;
  // generic accessor at fixed offsets (ii - i, jj - j in -1..2)
  MSCX_HD VKey operator()(int ii, int jj) const {
    const int c = ii - i + 1, r = jj - j + 1;
    const uint32_t f0 = r == 0 ? w00 : r == 1 ? w10 : r == 2 ? w20 : w30;
    const uint32_t f1 = r == 0 ? w01 : r == 1 ? w11 : r == 2 ? w21 : w31;
    const uint32_t f2 = r == 0 ? w02 : r == 1 ? w12 : r == 2 ? w22 : w32;
    const uint32_t f3 = r == 0 ? w03 : r == 1 ? w13 : r == 2 ? w23 : w33;
    return VKey{c == 0 ? f0 : c == 1 ? f1 : c == 2 ? f2 : f3, ii, jj};
  }
  // next row: rows move down by one, n0..n3 = keys of (i-1..i+2, j+3)
  MSCX_HD void slide(uint32_t n0, uint32_t n1, uint32_t n2, uint32_t n3) {
    w00 = w10, w01 = w11, w02 = w12, w03 = w13;
    w10 = w20, w11 = w21, w12 = w22, w13 = w23;
    w20 = w30, w21 = w31, w22 = w32, w23 = w33;
    w30 = n0, w31 = n1, w32 = n2, w33 = n3;
    ++j;
  }
};

// vertex_cand from the vertex key and its four neighbours (NX, NY before the
// vertex, PY, PX after it) and the usability of each direction.
MSCX_HD int vcand_keys(uint32_t fv, uint32_t fnx, uint32_t fny, uint32_t fpy, uint32_t fpx, bool anx, bool any,
                       bool apy, bool apx) {
  const uint32_t enx = (anx && fnx <= fv) ? fnx : KEY_INF;
  const uint32_t eny = (any && fny <= fv) ? fny : KEY_INF;
  const uint32_t epy = (apy && fpy < fv) ? fpy : KEY_INF;
  const uint32_t epx = (apx && fpx < fv) ? fpx : KEY_INF;
  // every value is computed before it is selected, so the compiler keeps
  // this (and the rules below) free of divergent branches
  const bool l1 = enx <= eny, l2 = epy <= epx;
  const uint32_t e1 = l1 ? enx : eny, e2 = l2 ? epy : epx;
  const bool l = e1 <= e2;
  const int d1 = l1 ? DIR_NX : DIR_NY, d2 = l2 ? DIR_PY : DIR_PX;
  const int d = l ? d1 : d2;
  return (e1 & e2) == KEY_INF ? DIR_NONE : d;  // no usable edge: all four are KEY_INF
}
// Edge candidates.  A quad is a usable cofacet of one of its edges exactly
// when that edge is its largest, i.e. joins its largest corner to the larger
// of that corner's two neighbours; equivalently, each end of the edge is
// above the corner diagonally across from it.  When both quads of an edge
// are usable, the end of the edge that is not the quad's largest corner is
// above its neighbour off the edge, so each quad's smallest corner is one of
// its two corners off the edge; those are distinct vertices, so quad_lt on
// the sorted corners is decided by the smaller of them alone.
//
// hedge_cand of the edge A=(x,y)-B=(x+1,y): Au, Bu the row above, Ad, Bd
// below.  Between equal values the key in column x, then the lower row, is
// the smaller.
MSCX_HD int hcand_keys(uint32_t fA, uint32_t fB, uint32_t fAu, uint32_t fBu, uint32_t fAd, uint32_t fBd) {
  const bool up = fAu <= fB && fBu < fA;
  const bool dn = fAd <= fB && fBd < fA;
  const bool cu = fAu <= fBu, cd = fAd <= fBd;  // smaller corner in column x
  const uint32_t mu = cu ? fAu : fBu, md = cd ? fAd : fBd;
  const bool up_first = mu < md || (mu == md && cu && !cd);
  const int d_up = (dn && !up_first) ? DIR_NY : DIR_PY, d_dn = dn ? DIR_NY : DIR_NONE;
  return up ? d_up : d_dn;
}
// vedge_cand of the edge A=(x,y)-B=(x,y+1): Ra, Rb the column right, La, Lb
// left.  Between equal values the key in the left column is the smaller.
MSCX_HD int ecand_keys(uint32_t fA, uint32_t fB, uint32_t fRa, uint32_t fRb, uint32_t fLa, uint32_t fLb) {
  const bool rt = fRa < fB && fRb < fA;
  const bool lf = fLa <= fB && fLb <= fA;
  const uint32_t mr = fRa < fRb ? fRa : fRb, ml = fLa < fLb ? fLa : fLb;
  const bool rt_first = mr < ml;
  const int d_rt = (lf && !rt_first) ? DIR_NX : DIR_PX, d_lf = lf ? DIR_NX : DIR_NONE;
  return rt ? d_rt : d_lf;
}

// Boundary classes of vertex column x / row y (INTERIOR: none applies).
struct VClass {
  bool in;     // x < vx
  bool lo;     // x > 0
  bool hi;     // x < vx - 1
  bool rv;     // vertex coordinate 2x inside the rect
  bool re;     // edge coordinate 2x+1 inside the rect
  bool rm;     // edge coordinate 2x-1 inside the rect
  bool sv;     // 2x is a boundary-cancellable side of the rect
  bool se;     // 2x+1 likewise
  bool sm;     // 2x-1 likewise
};
MSCX_HD VClass vclass(int x, int n, int r0, int r1, int cmax) {
  VClass c;
  c.in = x < n;
  c.lo = x > 0;
  c.hi = x < n - 1;
  c.rv = 2 * x >= r0 && 2 * x <= r1;
  c.re = 2 * x + 1 >= r0 && 2 * x + 1 <= r1;
  c.rm = 2 * x - 1 >= r0 && 2 * x - 1 <= r1;
  auto side = [&](int v) { return (v == r0 && r0 > 0) || (v == r1 && r1 < cmax); };
  c.sv = side(2 * x);
  c.se = side(2 * x + 1);
  c.sm = side(2 * x - 1);
  return c;
}

template <bool INTERIOR>
MSCX_HD int vcand_at(uint32_t fv, uint32_t fnx, uint32_t fny, uint32_t fpy, uint32_t fpx, const VClass &cx,
                     const VClass &cy) {
  if (INTERIOR) return vcand_keys(fv, fnx, fny, fpy, fpx, true, true, true, true);
  if (!(cx.in && cy.in && cx.rv && cy.rv)) return DIR_NONE;
  const bool colB = !cx.lo || !cx.hi, rowB = !cy.lo || !cy.hi;
  const bool hx = !colB || rowB, vy = !rowB || colB;
  return vcand_keys(fv, fnx, fny, fpy, fpx, cx.lo && hx, cy.lo && vy, cy.hi && vy, cx.hi && hx);
}

// Candidates at the window's centre (i, j) = (K.i, K.j): the vertex and
// the vertical edge (i, j)-(i, j+1).
template <bool INTERIOR>
MSCX_HD int vc_center(const KeyWin &K, const TileGeom &g) {
  const VClass c0 = vclass(K.i, g.vx, g.rx0, g.rx1, g.cw - 1), r0 = vclass(K.j, g.vy, g.ry0, g.ry1, g.ch - 1);
  return vcand_at<INTERIOR>(K.w11, K.w10, K.w01, K.w21, K.w12, c0, r0);
}
template <bool INTERIOR>
MSCX_HD int ec_center(const KeyWin &K, const TileGeom &g) {
  int e = ecand_keys(K.w11, K.w21, K.w12, K.w22, K.w10, K.w20);
  if (!INTERIOR) {
    const VClass c0 = vclass(K.i, g.vx, g.rx0, g.rx1, g.cw - 1);
    const VClass r0 = vclass(K.j, g.vy, g.ry0, g.ry1, g.ch - 1), r1 = vclass(K.j + 1, g.vy, g.ry0, g.ry1, g.ch - 1);
    if (!(c0.in && r1.in && c0.rv && r0.re && c0.lo && c0.hi)) e = DIR_NONE;
  }
  return e;
}

// ... and the horizontal edge (i, j)-(i+1, j).
template <bool INTERIOR>
MSCX_HD int hc_center(const KeyWin &K, const TileGeom &g) {
  int h = hcand_keys(K.w11, K.w12, K.w21, K.w22, K.w01, K.w02);
  if (!INTERIOR) {
    const VClass c0 = vclass(K.i, g.vx, g.rx0, g.rx1, g.cw - 1), c1 = vclass(K.i + 1, g.vx, g.rx0, g.rx1, g.cw - 1);
    const VClass r0 = vclass(K.j, g.vy, g.ry0, g.ry1, g.ch - 1);
    if (!(c1.in && r0.in && c0.re && r0.rv && r0.lo && r0.hi)) h = DIR_NONE;
  }
  return h;
}

// Candidates of the first row of a walk: vertices (i, j), (i+1, j) and the
// edge (i, j)-(i+1, j).
template <bool INTERIOR>
MSCX_HD void row_init(const KeyWin &K, const TileGeom &g, int &vc00, int &vc10, int &hc00) {
  const VClass c1 = vclass(K.i + 1, g.vx, g.rx0, g.rx1, g.cw - 1), r0 = vclass(K.j, g.vy, g.ry0, g.ry1, g.ch - 1);
  vc00 = vc_center<INTERIOR>(K, g);
  vc10 = vcand_at<INTERIOR>(K.w12, K.w11, K.w02, K.w22, K.w13, c1, r0);
  hc00 = hc_center<INTERIOR>(K, g);
}

// New candidates of a row step in the lane's own column: vertex (i, j+1),
// horizontal edge (i, j+1)-(i+1, j+1), vertical edge (i, j)-(i, j+1).
template <bool INTERIOR>
MSCX_HD void row_own(const KeyWin &K, const TileGeom &g, int &vc01, int &hc01, int &ec00) {
  const int i = K.i, j = K.j;
  const VClass c0 = vclass(i, g.vx, g.rx0, g.rx1, g.cw - 1), c1 = vclass(i + 1, g.vx, g.rx0, g.rx1, g.cw - 1);
  const VClass r0 = vclass(j, g.vy, g.ry0, g.ry1, g.ch - 1), r1 = vclass(j + 1, g.vy, g.ry0, g.ry1, g.ch - 1);
  vc01 = vcand_at<INTERIOR>(K.w21, K.w20, K.w11, K.w31, K.w22, c0, r1);
  int h = hcand_keys(K.w21, K.w22, K.w31, K.w32, K.w11, K.w12);
  int e = ecand_keys(K.w11, K.w21, K.w12, K.w22, K.w10, K.w20);
  if (!INTERIOR) {
    if (!(c1.in && r1.in && c0.re && r1.rv && r1.lo && r1.hi)) h = DIR_NONE;
    if (!(c0.in && r1.in && c0.rv && r0.re && c0.lo && c0.hi)) e = DIR_NONE;
  }
  hc01 = h;
  ec00 = e;
}

// The same step's candidates of the next column (vertex (i+1, j+1), edge
// (i+1, j)-(i+1, j+1)), which the kernel takes from the neighbouring lane.
template <bool INTERIOR>
MSCX_HD void row_next(const KeyWin &K, const TileGeom &g, int &vc11, int &ec10) {
  const int i = K.i, j = K.j;
  const VClass c1 = vclass(i + 1, g.vx, g.rx0, g.rx1, g.cw - 1);
  const VClass r0 = vclass(j, g.vy, g.ry0, g.ry1, g.ch - 1), r1 = vclass(j + 1, g.vy, g.ry0, g.ry1, g.ch - 1);
  vc11 = vcand_at<INTERIOR>(K.w22, K.w21, K.w12, K.w32, K.w23, c1, r1);
  int e = ecand_keys(K.w12, K.w22, K.w13, K.w23, K.w11, K.w21);
  if (!INTERIOR && !(c1.in && r1.in && c1.rv && r0.re && c1.lo && c1.hi)) e = DIR_NONE;
  ec10 = e;
}

// Pair directions of the cells (2i, 2j), (2i+1, 2j), (2i, 2j+1),
// (2i+1, 2j+1) (code_from_cands) from the seven candidates they depend on.
// An edge without a usable quad is paired with its higher end when that
// vertex's candidate is this edge: a vertex only picks an edge to a lower
// neighbour, so the candidate alone says which end is higher, and at most
// one end can pick the edge.  A quad is paired with its largest edge when
// that edge's candidate is the quad; an edge only picks a quad whose largest
// edge it is, so at most one of the four edges can point at the quad and
// the one that does is its largest.
MSCX_HD void row_cell_dirs(int vc00, int vc10, int hc00, int vc01, int hc01, int ec00, int ec10, int &dv, int &dh,
                           int &de, int &dq) {
  dv = vc00;
  const int hv = vc00 == DIR_PX ? DIR_NX : (vc10 == DIR_NX ? DIR_PX : DIR_NONE);
  dh = hc00 != DIR_NONE ? hc00 : hv;
  const int ev = vc00 == DIR_PY ? DIR_NY : (vc01 == DIR_NY ? DIR_PY : DIR_NONE);
  de = ec00 != DIR_NONE ? ec00 : ev;
  const int q1 = ec10 == DIR_NX ? DIR_PX : DIR_NONE;
  const int q2 = ec00 == DIR_PX ? DIR_NX : q1;
  const int q3 = hc01 == DIR_NY ? DIR_PY : q2;
  dq = hc00 == DIR_PY ? DIR_NY : q3;
}

// Code of a cell strictly inside the interior rect from its pair direction.
MSCX_HD uint8_t interior_code(int d) { return (uint8_t)((d ? CF_PAIRED : CF_CRITICAL) | (d << 2)); }

// The four codes of the same cells.  Cells outside the ext get code 0.
template <bool INTERIOR>
MSCX_HD void row_cell_codes(const KeyWin &K, const TileGeom &g, int vc00, int vc10, int hc00, int vc01, int hc01,
                            int ec00, int ec10, uint8_t &cv, uint8_t &chh, uint8_t &ce, uint8_t &cq) {
  const int i = K.i, j = K.j;
  const VClass c0 = vclass(i, g.vx, g.rx0, g.rx1, g.cw - 1), c1 = vclass(i + 1, g.vx, g.rx0, g.rx1, g.cw - 1);
  const VClass r0 = vclass(j, g.vy, g.ry0, g.ry1, g.ch - 1), r1 = vclass(j + 1, g.vy, g.ry0, g.ry1, g.ch - 1);
  int dv, dh, de, dq;
  row_cell_dirs(vc00, vc10, hc00, vc01, hc01, ec00, ec10, dv, dh, de, dq);
  if (INTERIOR) {
    cv = interior_code(dv);
    chh = interior_code(dh);
    ce = interior_code(de);
    cq = interior_code(dq);
  } else {
    // flags: in rect (exterior unpaired cells are not critical) and the
    // boundary-cancellable sites of a cell and of its pair
    auto site = [&](bool inx, bool iny, bool sx, bool sy) { return inx && iny && (sx || sy); };
    const bool in_v = c0.rv && r0.rv, in_h = c0.re && r0.rv, in_e = c0.rv && r0.re, in_q = c0.re && r0.re;
    const bool s_v = site(c0.rv, r0.rv, c0.sv, r0.sv), s_h = site(c0.re, r0.rv, c0.se, r0.sv);
    const bool s_e = site(c0.rv, r0.re, c0.sv, r0.se), s_q = site(c0.re, r0.re, c0.se, r0.se);
    // sites of the cells each of the four cells can be paired with
    const bool s_vpx = s_h;                                  // (2i+1, 2j)
    const bool s_vnx = site(c0.rm, r0.rv, c0.sm, r0.sv);     // (2i-1, 2j)
    const bool s_vpy = s_e;                                  // (2i, 2j+1)
    const bool s_vny = site(c0.rv, r0.rm, c0.sv, r0.sm);     // (2i, 2j-1)
    const bool s_hpx = site(c1.rv, r0.rv, c1.sv, r0.sv);     // (2i+2, 2j)
    const bool s_hnx = s_v;                                  // (2i, 2j)
    const bool s_hpy = s_q;                                  // (2i+1, 2j+1)
    const bool s_hny = site(c0.re, r0.rm, c0.se, r0.sm);     // (2i+1, 2j-1)
    const bool s_epx = s_q;                                  // (2i+1, 2j+1)
    const bool s_enx = site(c0.rm, r0.re, c0.sm, r0.se);     // (2i-1, 2j+1)
    const bool s_epy = site(c0.rv, r1.rv, c0.sv, r1.sv);     // (2i, 2j+2)
    const bool s_eny = s_v;                                  // (2i, 2j)
    const bool s_qpx = site(c1.rv, r0.re, c1.sv, r0.se);     // (2i+2, 2j+1)
    const bool s_qnx = s_e;                                  // (2i, 2j+1)
    const bool s_qpy = site(c0.re, r1.rv, c0.se, r1.sv);     // (2i+1, 2j+2)
    const bool s_qny = s_h;                                  // (2i+1, 2j)
    auto pick = [](int d, bool px, bool nx, bool py, bool ny) {
      return d == DIR_PX ? px : d == DIR_NX ? nx : d == DIR_PY ? py : d == DIR_NY ? ny : false;
    };
    auto mk = [](int d, bool inr, bool self_site, bool pair_site) -> uint8_t {
      if (d == DIR_NONE) return (uint8_t)(inr ? CF_CRITICAL : 0);
      return (uint8_t)(CF_PAIRED | ((self_site || pair_site) ? CF_CRITICAL : 0) | (d << 2));
    };
    cv = mk(dv, in_v, s_v, pick(dv, s_vpx, s_vnx, s_vpy, s_vny));
    chh = mk(dh, in_h, s_h, pick(dh, s_hpx, s_hnx, s_hpy, s_hny));
    ce = mk(de, in_e, s_e, pick(de, s_epx, s_enx, s_epy, s_eny));
    cq = mk(dq, in_q, s_q, pick(dq, s_qpx, s_qnx, s_qpy, s_qny));
    if (!(c0.in && r0.in)) cv = 0;
    if (!(c1.in && r0.in)) chh = 0;
    if (!(c0.in && r1.in)) ce = 0;
    if (!(c1.in && r1.in)) cq = 0;
  }
}

// One full row step as the CPU checks run it: own and next-column
// candidates, the four codes, and the carried candidates advanced to row j+1.
template <bool INTERIOR>
MSCX_HD void row_codes(const KeyWin &K, const TileGeom &g, int &vc00, int &vc10, int &hc00, uint8_t &cv, uint8_t &chh,
                       uint8_t &ce, uint8_t &cq) {
  int vc01, hc01, ec00, vc11, ec10;
  row_own<INTERIOR>(K, g, vc01, hc01, ec00);
  row_next<INTERIOR>(K, g, vc11, ec10);
  row_cell_codes<INTERIOR>(K, g, vc00, vc10, hc00, vc01, hc01, ec00, ec10, cv, chh, ce, cq);
  vc00 = vc01;
  vc10 = vc11;
  hc00 = hc01;
}

}  // namespace mscx
