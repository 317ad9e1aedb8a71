// Test-only CPU build of the per-cell gradient rules (cell_rules.hpp), so
// the rules the HIP kernel runs can be checked against the oracle without a
// GPU.  Not linked into libmscx.so.
#include <cstddef>
#include <cstdint>
#include <type_traits>
#include "../mscomplex_amd/csrc/cell_rules.hpp"

extern "C" int cr_codes(const float *win, const int *rect, const int *ext, uint8_t *codes) {
  mscx::TileGeom g;
  g.vx = (ext[2] - ext[0]) / 2 + 1;
  g.vy = (ext[3] - ext[1]) / 2 + 1;
  g.cw = 2 * g.vx - 1;
  g.ch = 2 * g.vy - 1;
  g.rx0 = rect[0] - ext[0];
  g.ry0 = rect[1] - ext[1];
  g.rx1 = rect[2] - ext[0];
  g.ry1 = rect[3] - ext[1];
  auto K = [&](int i, int j) -> uint64_t {
    if (i < -1 || j < -1 || i > g.vx || j > g.vy) throw 1;  // beyond the 1-vertex ring: a bug
    if (i < 0 || j < 0 || i >= g.vx || j >= g.vy) return 0x12345ull;  // read and discarded
    return mscx::vkey(mscx::fkey(win[(size_t)j * g.vx + i]), i, j);
  };
  try {
    for (int y = 0; y < g.ch; ++y)
      for (int x = 0; x < g.cw; ++x) codes[(size_t)y * g.cw + x] = mscx::cell_code(K, g, x, y);
  } catch (int) {
    return -1;
  }
  // the decomposed form the kernels run, with the split (value, x, y) keys
  // they use, must agree with the per-cell form on packed 64-bit keys
  // the split-key rules read neighbours just outside the ext (and discard
  // them), as the kernels do from their register window
  auto KS = [&](int i, int j) -> mscx::VKey {
    if (i < -1 || j < -1 || i > g.vx || j > g.vy) throw 1;
    if (i < 0 || j < 0 || i >= g.vx || j >= g.vy) return mscx::VKey{0x12345u, i, j};
    return mscx::VKey{mscx::fkey(win[(size_t)j * g.vx + i]), i, j};
  };
  auto VC = [&](int i, int j) {
    return mscx::in_rect(g, 2 * i, 2 * j) ? mscx::vertex_cand(KS, g, i, j) : mscx::DIR_NONE;
  };
  auto HC = [&](int i, int j) {
    return mscx::in_rect(g, 2 * i + 1, 2 * j) ? mscx::hedge_cand(KS, g, i, j) : mscx::DIR_NONE;
  };
  auto EC = [&](int i, int j) {
    return mscx::in_rect(g, 2 * i, 2 * j + 1) ? mscx::vedge_cand(KS, g, i, j) : mscx::DIR_NONE;
  };
  try {
    for (int y = 0; y < g.ch; ++y)
      for (int x = 0; x < g.cw; ++x)
        if (mscx::code_from_cands(KS, VC, HC, EC, g, x, y) != codes[(size_t)y * g.cw + x]) return -2;
  } catch (int) {
    return -3;
  }
  // the kernels' interior fast path (no boundary tests) on every cell whose
  // neighbourhood lies strictly inside the interior rect
  auto VI = [&](int i, int j) { return mscx::vertex_cand<true>(KS, g, i, j); };
  auto HI = [&](int i, int j) { return mscx::hedge_cand<true>(KS, g, i, j); };
  auto EI = [&](int i, int j) { return mscx::vedge_cand<true>(KS, g, i, j); };
  try {
    for (int y = 0; y < g.ch; ++y)
      for (int x = 0; x < g.cw; ++x) {
        const int i = x >> 1, j = y >> 1;
        if (!(2 * i - 1 > g.rx0 && 2 * (i + 1) + 1 < g.rx1 && 2 * j - 1 > g.ry0 && 2 * (j + 1) + 1 < g.ry1)) continue;
        if (mscx::code_from_cands<true>(KS, VI, HI, EI, g, x, y) != codes[(size_t)y * g.cw + x]) return -4;
      }
  } catch (int) {
    return -5;
  }
  // the kernels' row walk (row_init / row_codes): every column walked from
  // row 0 with the generic form, and with the interior form over each run
  // of rows whose cells and window lie strictly inside the interior rect
  auto fk = [&](int i, int j) -> uint32_t {
    if (i < 0 || j < 0 || i >= g.vx || j >= g.vy) return 0x12345u;
    return mscx::fkey(win[(size_t)j * g.vx + i]);
  };
  auto load = [&](int i, int j) {
    mscx::KeyWin K;
    uint32_t *w = &K.w00;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) w[4 * r + c] = fk(i - 1 + c, j - 1 + r);
    K.i = i;
    K.j = j;
    return K;
  };
  auto code_at = [&](int x, int y) -> int {
    if (x >= g.cw || y >= g.ch) return 0;
    return codes[(size_t)y * g.cw + x];
  };
  auto walk = [&](auto interior, int i, int j0, int j1) -> bool {
    constexpr bool IN = decltype(interior)::value;
    mscx::KeyWin K = load(i, j0);
    int vc00, vc10, hc00;
    mscx::row_init<IN>(K, g, vc00, vc10, hc00);
    for (int j = j0; j < j1; ++j) {
      uint8_t cv, chh, ce, cq;
      mscx::row_codes<IN>(K, g, vc00, vc10, hc00, cv, chh, ce, cq);
      if (cv != code_at(2 * i, 2 * j) || chh != code_at(2 * i + 1, 2 * j) || ce != code_at(2 * i, 2 * j + 1) ||
          cq != code_at(2 * i + 1, 2 * j + 1))
        return false;
      K.slide(fk(i - 1, j + 3), fk(i, j + 3), fk(i + 1, j + 3), fk(i + 2, j + 3));
    }
    return true;
  };
  for (int i = 0; i < g.vx; ++i) {
    if (!walk(std::integral_constant<bool, false>(), i, 0, g.vy)) return -6;
    if (!(2 * i - 1 > g.rx0 && 2 * (i + 1) + 1 < g.rx1)) continue;
    int j = 0;
    while (j < g.vy) {
      while (j < g.vy && !(2 * j - 1 > g.ry0 && 2 * (j + 1) + 1 < g.ry1)) ++j;
      int j1 = j;
      while (j1 < g.vy && (2 * j1 - 1 > g.ry0 && 2 * (j1 + 1) + 1 < g.ry1)) ++j1;
      if (j < j1 && !walk(std::integral_constant<bool, true>(), i, j, j1)) return -7;
      j = j1;
    }
  }
  return 0;
}

// Codes of the cells [x0, x0+w) x [y0, y0+h) of a tile (ext-local cell
// coordinates) from the per-cell rules, reading the tile's whole window.
extern "C" int cr_codes_region(const float *win, const int *rect, const int *ext, int x0, int y0, int w, int h,
                               uint8_t *codes) {
  mscx::TileGeom g;
  g.vx = (ext[2] - ext[0]) / 2 + 1;
  g.vy = (ext[3] - ext[1]) / 2 + 1;
  g.cw = 2 * g.vx - 1;
  g.ch = 2 * g.vy - 1;
  g.rx0 = rect[0] - ext[0];
  g.ry0 = rect[1] - ext[1];
  g.rx1 = rect[2] - ext[0];
  g.ry1 = rect[3] - ext[1];
  auto K = [&](int i, int j) -> uint64_t {
    if (i < -1 || j < -1 || i > g.vx || j > g.vy) throw 1;
    if (i < 0 || j < 0 || i >= g.vx || j >= g.vy) return 0x12345ull;
    return mscx::vkey(mscx::fkey(win[(size_t)j * g.vx + i]), i, j);
  };
  try {
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) codes[(size_t)y * w + x] = mscx::cell_code(K, g, x0 + x, y0 + y);
  } catch (int) {
    return -1;
  }
  return 0;
}

// The row-step edge candidates (hcand_keys / ecand_keys, which decide an
// edge with both quads usable by the smaller off-edge corners) against the
// per-cell rules (hedge_cand / vedge_cand with quad_lt on sorted corners),
// over every assignment of nv values to the six vertices around a
// horizontal and a vertical edge.  Returns the number of disagreements.
extern "C" long cr_edge_forms_check(int nv) {
  mscx::TileGeom g;
  g.vx = g.vy = 4;
  g.cw = g.ch = 7;
  g.rx0 = g.ry0 = 0;
  g.rx1 = g.ry1 = 6;
  float win[16];
  auto KS = [&](int i, int j) { return mscx::VKey{mscx::fkey(win[4 * j + i]), i, j}; };
  auto fk = [&](int i, int j) { return mscx::fkey(win[4 * j + i]); };
  long total = 1;
  for (int k = 0; k < 6; ++k) total *= nv;
  long bad = 0;
  for (long code = 0; code < total; ++code) {
    int v[6];
    long c = code;
    for (int k = 0; k < 6; ++k) v[k] = (int)(c % nv), c /= nv;
    // horizontal edge (1,1)-(2,1): rows 0..2 of columns 1, 2
    for (float &w : win) w = 0.5f;
    win[4 * 1 + 1] = (float)v[0], win[4 * 1 + 2] = (float)v[1];
    win[4 * 2 + 1] = (float)v[2], win[4 * 2 + 2] = (float)v[3];
    win[4 * 0 + 1] = (float)v[4], win[4 * 0 + 2] = (float)v[5];
    if (mscx::hedge_cand(KS, g, 1, 1) != mscx::hcand_keys(fk(1, 1), fk(2, 1), fk(1, 2), fk(2, 2), fk(1, 0), fk(2, 0)))
      ++bad;
    // vertical edge (1,1)-(1,2): columns 0..2 of rows 1, 2
    for (float &w : win) w = 0.5f;
    win[4 * 1 + 1] = (float)v[0], win[4 * 2 + 1] = (float)v[1];
    win[4 * 1 + 2] = (float)v[2], win[4 * 2 + 2] = (float)v[3];
    win[4 * 1 + 0] = (float)v[4], win[4 * 2 + 0] = (float)v[5];
    if (mscx::vedge_cand(KS, g, 1, 1) != mscx::ecand_keys(fk(1, 1), fk(1, 2), fk(2, 1), fk(2, 2), fk(0, 1), fk(0, 2)))
      ++bad;
  }
  return bad;
}
