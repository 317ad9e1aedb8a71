// mscx_tile: one GridDataset (rect + ext window) resident in HBM.
#pragma once

#include <memory>
#include <vector>

#include "kernels.hpp"
#include "runtime.hpp"

struct mscx_tile {
  int32_t rect[4];
  int32_t ext[4];
  mscx::TileDesc d{};     // device pointers and geometry
  void *fixed = nullptr;  // one allocation: codes, counts, offsets, trace pointers, owners
  size_t fixed_bytes = 0;
  float *verts_owned = nullptr;
  size_t verts_cap = 0;
  // crit, fn, conn (x4), pair of the last work: one block shared by all
  // tiles of that launch (one device->host copy serves them all)
  std::shared_ptr<void> crit_block;
  size_t crit_off = 0;  // byte offset of this tile's arrays in the block
  size_t crit_block_bytes = 0;
  uint32_t K = 0;
  bool has_input = false;
  bool worked = false;
};

namespace mscx {

// Runs gradient, collation, trace, owners and incidence for n tiles with one
// launch per stage, then synchronises.  Sets K / worked on every tile.
void run_tiles(mscx_tile *const *tiles, int n);

// Block tile height (GRAD_TY or GRAD_TY_LOW rows) a tile of vy vertex rows
// takes, from the table mscx_set_block_rows installs.
int block_rows_for(int vy);

// Size of the device allocation a tile over (rect, ext) makes at creation.
size_t tile_fixed_bytes(const int32_t rect[4], const int32_t ext[4]);
// Throws MSCX_EINVAL unless (rect, ext) is a valid tile.
void check_rects(const int32_t *rect, const int32_t *ext);
// Geometry of a tile over ext in d and the byte offsets of its device
// arrays in one block: o[0] codes, o[1] segment offsets, o[2] scan
// partials, o[3] critical bits, o[4] owners, o[5] local terminals, o[6]
// ring words (ringidx follows them); returns the block size.
size_t fixed_layout(const int32_t rect[4], const int32_t ext[4], TileDesc &d, size_t o[7]);

// Byte offsets of the crit / fn / conn / pair arrays for K critical cells.
struct CritLayout {
  size_t fn, conn, pair, total;
  explicit CritLayout(size_t K);
};

// Host copies of a tile's critical data after work (K entries each).
void tile_critical(const mscx_tile *t, std::vector<uint32_t> &crit, std::vector<float> &fn,
                   std::vector<uint32_t> &conn4, std::vector<uint32_t> &pair);
// Host copies of a tile's codes and owners over ext, x fastest (unpitched).
void tile_codes(const mscx_tile *t, std::vector<uint8_t> &codes);
// Owners as packed cell ids (the reference's m_cell_own) or as indices into
// the tile's critical list (the device representation).
void tile_owners(const mscx_tile *t, std::vector<uint32_t> &own);
void tile_owner_index(const mscx_tile *t, std::vector<uint32_t> &own);

}  // namespace mscx
