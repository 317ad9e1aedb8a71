// Per-tile device descriptor shared by the host launcher and the kernels.
// Every launch processes an array of tiles; block b of a launch belongs to
// tile t when blk_prefix[t] <= b < blk_prefix[t+1].
#pragma once

#include <stdint.h>

#include "cell_rules.hpp"

namespace mscx {

constexpr int GRAD_TX = 64;       // vertices per gradient/trace block tile in x (one per lane)
constexpr int GRAD_TY = 32;       // vertex rows per block tile (the default shape)
constexpr int GRAD_TY_LOW = 16;   // the other block tile height the kernels carry
constexpr int SEG = 2 * GRAD_TX;  // cells per collation segment: one block tile row
constexpr int SCAN_CHUNK = 2048;  // counts per scan block

// Perimeter ("ring") of a block tile of ty rows, per forest (vertices, then
// quads at ring_q(ty)): four side arrays, the bottom row (lj = 0) at
// [0, TX), the top row (lj = ty-1) at [TX, 2TX), the left column (li = 0) at
// [2TX, 2TX+ty) and the right column (li = TX-1) at [2TX+ty, 2TX+2ty); a
// corner has a word in two arrays, both with the same value.  A gradient
// path that leaves a block through side s at position p enters its
// neighbour on the array of the opposite side, at ring_entry_base(s, ty) + p.
MSCX_HD constexpr int ring_q(int ty) { return 2 * GRAD_TX + 2 * ty; }  // ring words per forest
MSCX_HD constexpr int ring_w(int ty) { return 2 * ring_q(ty); }        // ring words per block

// side arrays entered through side s = 0 (+x), 1 (-x), 2 (+y), 3 (-y): the
// neighbour's left column, right column, bottom row, top row
MSCX_HD int ring_entry_base(int s, int ty) {
  return s == 0 ? 2 * GRAD_TX : s == 1 ? 2 * GRAD_TX + ty : s == 2 ? 0 : GRAD_TX;
}
MSCX_HD void ring_at(int p, int ty, int &li, int &lj) {
  if (p < GRAD_TX) {
    li = p;
    lj = 0;
  } else if (p < 2 * GRAD_TX) {
    li = p - GRAD_TX;
    lj = ty - 1;
  } else if (p < 2 * GRAD_TX + ty) {
    li = 0;
    lj = p - 2 * GRAD_TX;
  } else {
    li = GRAD_TX - 1;
    lj = p - (2 * GRAD_TX + ty);
  }
}

// Block tiles of a tile whose vertex window is g, with nseg block columns,
// at ty rows per block; and the ring words its allocation holds (enough for
// either block height, so the height can be chosen after allocation).
MSCX_HD int64_t tile_blocks(const TileGeom &g, int nseg, int ty) {
  return (int64_t)nseg * ((g.vy + ty - 1) / ty);
}
MSCX_HD int64_t ring_words_max(const TileGeom &g, int nseg) {
  const int64_t a = tile_blocks(g, nseg, GRAD_TY) * ring_w(GRAD_TY);
  const int64_t b = tile_blocks(g, nseg, GRAD_TY_LOW) * ring_w(GRAD_TY_LOW);
  return a > b ? a : b;
}

// Local terminal of a vertex or quad entry after the block-local trace
// (16 bits; in LDS an entry with bit 15 clear is a link: the byte offset of
// another entry of the block):
//   0x8000 | r<<7 | k   root: the critical cell in cell row r (0..63) of the
//                       block, k-th critical cell of that row's segment
//   0xA000 | s<<6 | p   exit: the path leaves through side s (0 +x, 1 -x,
//                       2 +y, 3 -y) at position p along that side
//   0xC000              none: the cell has no owner
constexpr uint32_t T_ROOT = 0x8000u, T_EXIT = 0xA000u, T_NONE = 0xC000u;
MSCX_HD uint32_t term_kind(uint32_t t) { return (t >> 13) & 3u; }  // 0 root, 1 exit, 2 none

struct TileDesc {
  // geometry (local coordinates relative to the ext rect's bottom-left cell)
  TileGeom g;
  int ox, oy;  // global cell coordinates of the ext rect's bottom-left
  int nseg;    // collation segments (= block tiles) per cell row (pitch / SEG)
  int pitch;   // cells per code / owner row (nseg * SEG)
  // input: ext window of vertex values, x fastest, row pitch in floats
  const float *verts;
  int64_t vpitch;
  // gradient codes, ch rows of `pitch` bytes
  uint8_t *codes;
  // collation: per (cell row, segment) the critical flags as two 64-bit
  // words (cells 2l and 2l+1 of the segment are bit l of word 0 and word 1)
  // and the exclusive scan of their counts
  uint32_t *seg_offs;    // ch * nseg
  uint64_t *critbits;    // ch * nseg * 2
  uint32_t *scan_part;   // per scan block partial sums
  uint32_t *nkrit;       // [0] = number of critical cells K
  // local terminals: per vertex position, vertex terminal | quad terminal << 16
  uint32_t *loc;
  int loc_pitch;  // words per vertex row (nseg * GRAD_TX)
  // block tile height (GRAD_TY or GRAD_TY_LOW vertex rows)
  int ty;
  // ring words per block (ring_w(ty) each): after the gradient kernel a root
  // (ext-local cell row << 15 | segment << 7 | rank in the segment), OWN_NONE
  // or RING_MARK | ring index of the entry the path continues at; ringidx
  // holds the resolved critical index
  uint32_t *ring;
  uint32_t *ringidx;
  // critical list (raster order), their function values, saddle incidence
  // (facet owners then cofacet owners, critical indices or ~0) and pair index
  uint32_t *crit;
  float *crit_fn;
  uint32_t *crit_conn;  // 4 per critical cell
  uint32_t *crit_pair;  // critical index of the pair (boundary-cancellable) or ~0
  // owners: critical index of each cell's owner (the reference's m_cell_own
  // as an index into the critical list), OWN_NONE when unreached; ch rows of
  // own_pitch words
  uint32_t *own;
  int own_pitch;
  uint32_t kcap;  // capacity of the crit arrays
};

constexpr uint32_t OWN_NONE = 0xFFFFFFFFu;   // (-1,-1) / no owner
constexpr uint32_t RING_MARK = 0x80000000u;  // root words never have bit 31 set

// Critical index of ext-local cell (x, y), known to be critical: the scanned
// offset of its (row, segment) plus the critical cells before it there.
MSCX_HD uint32_t crit_rank(uint32_t base, uint64_t w0, uint64_t w1, int x) {
  const int l = (x & (SEG - 1)) >> 1;
  const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
  uint32_t r = (uint32_t)(__builtin_popcountll(w0 & below) + __builtin_popcountll(w1 & below));
  if (x & 1) r += (uint32_t)((w0 >> l) & 1u);
  return base + r;
}

// Per-leaf inputs of the manifold label pass (postMergeFillDiscs,
// grid_dataset.cpp:1142-1187): the owner index image and codes, and per
// critical index the label of a cell it owns (lab_owned: the cell is not
// critical) or of the critical cell itself (lab_self), as computed on the
// host from the owner redirects and the surviving extremum maps.
struct DiscDesc {
  const uint32_t *own;
  int64_t own_pitch;
  const uint8_t *codes;
  int64_t pitch;
  int ox, oy;              // ext bottom-left, global cell coords
  int rx0, ry0, rx1, ry1;  // interior rect, global cell coords
  const uint32_t *lab_owned;
  const uint32_t *lab_self;
  uint32_t nlab;
  uint32_t *lab_v;  // per interior vertex, raster order: critical-point index or ~0
  uint32_t *lab_q;  // per interior quad
  uint32_t *err;    // set when an owner is not a surviving extremum
  // grouping: the leaf's first sort key (labels are offset by it) and its
  // first element (its vertices, then its quads) in the sort input
  uint32_t lab_base;
  uint64_t elem_base;
};

constexpr uint32_t LAB_ERR = 0xFFFFFFFEu;  // label table entry: owner is no surviving extremum

// Per-leaf inputs of the disc tracing (postMergeFillDiscs' saddle
// 1-manifolds, grid_dataset.cpp:1189-1228): one thread per disc slot
// 2 i + j of the leaf's complex (j = 0 ascending, 1 descending).  A slot's
// list starts with its initial entries (init_cells[init_off[s] ..
// init_off[s+1])); a traced slot (bit s of `traced` set, a surviving saddle)
// grows by the successors of every entry in list order.
struct DiscTraceDesc {
  const uint8_t *codes;
  int64_t pitch;
  int ox, oy, ex1, ey1;    // ext, global cell coords
  int rx0, ry0, rx1, ry1;  // interior rect, global cell coords
  const uint32_t *init_off;    // nslot + 1
  const uint32_t *init_cells;
  const uint8_t *traced;       // nslot
  uint32_t nslot;
  uint32_t *count;             // nslot: list lengths (count pass; see DISC_TRACE_*)
  const uint32_t *out_off;     // nslot: list offsets in out (write pass)
  uint32_t *out;               // the lists (write pass)
  uint32_t *err;  // 2: a chain cell with two successors, an unpaired successor or a chain
                  // longer than the ext; 4: an unpaired head (not a gradient's V-paths)
};
constexpr int DISC_TRACE_CHAINS = 16;  // chains a thread follows for one disc
// A slot whose initial entries grow more chains than that is traced by a
// whole wave (a lane per chain, up to DISC_TRACE_WIDE_CAP chains): the
// thread's count pass marks it DISC_TRACE_WIDE, the wave's count pass
// replaces the mark by DISC_TRACE_WIDE_LEN | length, or by DISC_TRACE_HOST
// when even the wave's chain table is too small (the host traces that slot
// and the write passes leave its range alone).
constexpr int DISC_TRACE_WIDE_CAP = 1024;
constexpr uint32_t DISC_TRACE_WIDE_LEN = 0x80000000u;
constexpr uint32_t DISC_TRACE_WIDE = 0xFFFFFFFEu;
constexpr uint32_t DISC_TRACE_HOST = 0xFFFFFFFFu;

}  // namespace mscx
