// Test-only CPU build of the per-cell gradient rules (cell_rules.hpp), so
// the rules the HIP kernel runs can be checked against the oracle without a
// GPU.  Not linked into libmscx.so.
#include <cstddef>
#include <cstdint>
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
    if (i < 0 || j < 0 || i >= g.vx || j >= g.vy) throw 1;  // out-of-window access is a bug
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
  auto KS = [&](int i, int j) -> mscx::VKey {
    if (i < 0 || j < 0 || i >= g.vx || j >= g.vy) throw 1;
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
  return 0;
}
