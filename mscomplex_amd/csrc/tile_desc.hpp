// Per-tile device descriptor shared by the host launcher and the kernels.
// Every launch processes an array of tiles; block b of a launch belongs to
// tile t when blk_prefix[t] <= b < blk_prefix[t+1].
#pragma once

#include <stdint.h>

#include "cell_rules.hpp"

namespace mscx {

constexpr int SEG = 256;          // cells per collation segment (one wavefront row of the gradient kernel)
constexpr int GRAD_TX = 128;      // vertices per gradient tile in x (64 lanes x 2)
constexpr int GRAD_TY = 4;        // vertex rows per gradient tile (one per wave)
constexpr int SCAN_CHUNK = 2048;  // counts per scan block

struct TileDesc {
  // geometry (local coordinates relative to the ext rect's bottom-left cell)
  TileGeom g;
  int ox, oy;  // global cell coordinates of the ext rect's bottom-left
  int nseg;    // collation segments per cell row (pitch / SEG)
  int pitch;   // bytes per code row (nseg * SEG)
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
  // trace: per vertex / per quad successor; >=0 another vertex/quad,
  // -1 unowned, <= -2 resolved to critical index (-v - 2)
  int32_t *vptr;  // vy * vx
  int32_t *qptr;  // (vy - 1) * (vx - 1)
  // owners: packed cellid per cell (x | y << 16), ch rows of own_pitch words
  uint32_t *own;
  int own_pitch;
  uint32_t kcap;  // capacity of the crit arrays
};

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
