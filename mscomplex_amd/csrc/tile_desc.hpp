// Per-tile device descriptor shared by the host launcher and the kernels.
// Every launch processes an array of tiles; block b of a launch belongs to
// tile t when blk_prefix[t] <= b < blk_prefix[t+1].
#pragma once

#include <stdint.h>

#include "cell_rules.hpp"

namespace mscx {

constexpr int GRAD_TX = 64;       // vertices per gradient/trace block tile in x (one per lane)
constexpr int GRAD_TY = 32;       // vertex rows per block tile
constexpr int SEG = 2 * GRAD_TX;  // cells per collation segment: one block tile row
constexpr int SCAN_CHUNK = 2048;  // counts per scan block

struct TileDesc {
  // geometry (local coordinates relative to the ext rect's bottom-left cell)
  TileGeom g;
  int ox, oy;  // global cell coordinates of the ext rect's bottom-left
  int nseg;    // collation segments per cell row (pitch / SEG)
  int pitch;   // cells per code / owner row (nseg * SEG)
  // input: ext window of vertex values, x fastest, row pitch in floats
  const float *verts;
  int64_t vpitch;
  // gradient codes, ch rows of `pitch` bytes
  uint8_t *codes;
  // collation: per (cell row, segment) critical counts and their exclusive scan
  uint32_t *seg_counts;  // ch * nseg
  uint32_t *seg_offs;    // ch * nseg
  uint32_t *scan_part;   // per scan block partial sums
  uint32_t *nkrit;       // [0] = number of critical cells K
  // critical list (raster order), their function values, saddle incidence
  // (facet owners then cofacet owners, critical indices or ~0) and pair index
  uint32_t *crit;
  float *crit_fn;
  uint32_t *crit_conn;  // 4 per critical cell
  uint32_t *crit_pair;  // critical index of the pair (boundary-cancellable) or ~0
  // owners: packed cellid per cell (x | y << 16), ch rows of own_pitch
  // words; between the trace kernels an entry may hold OWN_MARK | packed id
  // of a cell whose owner it takes (a path that left its block tile)
  uint32_t *own;
  int own_pitch;
  // owner entries that hold a marker after the block-local trace: block b of
  // the gradient launch lists its entries in exits[b * 4*GRAD_TX*GRAD_TY ...]
  // and their count in nexit[b]
  uint32_t *exits;
  uint32_t *nexit;
  uint32_t kcap;  // capacity of the crit arrays
};

constexpr uint32_t OWN_NONE = 0xFFFFFFFFu;  // (-1,-1)
constexpr uint32_t OWN_MARK = 0x80000000u;  // cell ids never have bit 31 set (y <= 32766)
MSCX_HD bool own_is_mark(uint32_t w) { return (w & OWN_MARK) && w != OWN_NONE; }

// Per-leaf inputs of the manifold label pass (postMergeFillDiscs,
// grid_dataset.cpp:1142-1187): owner image, owner redirects of cancelled
// pairs, and the surviving extremum maps; tables are (cell, value) pairs
// sorted by packed cell id.
struct DiscDesc {
  const uint32_t *own;
  int64_t own_pitch;
  int ox, oy;              // ext bottom-left, global cell coords
  int rx0, ry0, rx1, ry1;  // interior rect, global cell coords
  const uint32_t *redir;
  uint32_t nredir;
  const uint32_t *minmap;
  uint32_t nmin;
  const uint32_t *maxmap;
  uint32_t nmax;
  uint32_t *lab_v;  // per interior vertex, raster order: critical-point index or ~0
  uint32_t *lab_q;  // per interior quad
  uint32_t *err;    // set when an owner is not a surviving extremum
};

}  // namespace mscx
