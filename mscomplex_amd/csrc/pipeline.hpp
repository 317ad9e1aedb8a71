// The GridDataManager compute order (grid_datamanager.cpp:596-662) as stages
// over sets of grids, shared by mscx_grid_run / mscx_batch_run
// (pipeline.cpp) and the serving loop (server.cpp).
#pragma once

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "complex.hpp"
#include "dmerge.hpp"
#include "runtime.hpp"
#include "tile.hpp"

namespace mscx {

struct LeafGeom {
  Rect rect, ext;
};

// Leaf (rect, ext) pairs of a grid's tiling (createPieces_quadtree).
void create_pieces(Rect r, Rect e, int level, int lx, int ly, bool two_d, std::vector<LeafGeom> &out);

}  // namespace mscx

// One GridDataManager run over one grid.
struct mscx_grid {
  int32_t sx = 0, sy = 0;
  double threshold = 0;
  uint32_t flags = 0;
  double work_limit = 0;  // simplify / un_simplify work limit (0: none; served requests set it)
  // Out-of-core mode (MSCX_RUN_OUT_OF_CORE): a piece that no phase is using
  // is kept as its dump words (or, with MSCX_SPILL_DIR set, as an archive
  // file there) instead of a live complex, and brought back when a phase or
  // a caller needs it; pieces[i] is then NULL.
  bool ooc = false;
  std::vector<std::unique_ptr<mscx::ComplexSpill>> spilled;
  std::string spill_dir;
  std::vector<char> spilled_file;
  std::mutex spill_mu;  // lazy restores by the accessors
  std::vector<mscx::LeafGeom> leaves;
  std::vector<mscx_complex *> pieces;  // leaves, then internal nodes in merge order; root last
  // Device merge tree (MSCX_RUN_DEVICE_MERGE): the pieces' complexes live in
  // HBM (dp[i]); pieces[i] is their host copy, made when a phase or a caller
  // needs it.
  bool dmerge = false;
  std::vector<std::unique_ptr<mscx::DComplex>> dp;
  std::vector<mscx_tile *> tiles;
  bool owns_tiles = true;         // false when the tiles belong to a resident grid of the server
  std::shared_ptr<void> keep;     // keeps the tiles' owner alive (server results)
  void *dev_input = nullptr;
  double t[5] = {0, 0, 0, 0, 0};  // compute, merge-up, simplify, merge-down, discs (seconds)
  ~mscx_grid();
};

namespace mscx {

double now_s();

// Validates arguments, uploads host input (unless MSCX_RUN_DEVICE_INPUT),
// creates and binds the leaf tiles.
void prepare_grid(mscx_grid *G, const float *data, int32_t sx, int32_t sy, int32_t levels, int32_t lx, int32_t ly,
                  double threshold, uint32_t flags);
// GPU stages of every leaf of every grid (one launch per stage) and the
// leaf complexes (writeout_connectivity).
void grids_compute(std::vector<mscx_grid *> &grids);
// mergePiecesUp; then simplification at G->threshold and mergePiecesDown.
void grid_merge_up(mscx_grid *G, bool inner_parallel);
void grid_simplify_down(mscx_grid *G, bool inner_parallel);
// postMergeFillDiscs for every leaf of the given grids (label pass in one launch).
void grids_discs(std::vector<mscx_grid *> &grids, CodesStage *pre = nullptr);
// HBM a grid's tiles and input take (for admission against a budget);
// the critical-point block is estimated at one critical cell in four.
size_t grid_device_bytes(int32_t sx, int32_t sy, int32_t levels, int32_t lx, int32_t ly, bool with_input);

// A raw grid file opened for the file-driven run (reader.cpp).
struct GridFile {
  int fd = -1;
  size_t bytes = 0;  // size_x * size_y * 4, the part that is read
  GridFile() = default;
  GridFile(const GridFile &) = delete;
  GridFile &operator=(const GridFile &) = delete;
  ~GridFile();
};
// Opens path and checks that it holds size_x * size_y floats (MSCX_EIO).
void open_grid_file(const char *path, int32_t sx, int32_t sy, GridFile &f);
// Streams the grid into dev (size_x * size_y floats) through two pinned
// slabs, each copy overlapping the next read; rejects NaN (MSCX_EINVAL).
void stream_grid_file(GridFile &f, int32_t sx, float *dev);

// Deep copy of a grid's pieces (the complexes, in HBM for a device merge
// tree); tiles are not copied.
void copy_pieces(const mscx_grid *src, mscx_grid *dst);
// HBM held by a grid's device-resident complexes.
size_t device_tree_bytes(const mscx_grid *G);
// The run flags with the merge tree's placement decided for an sx x sy grid
// (MSCX_RUN_DEVICE_MERGE from MSCX_DEVICE_MERGE_MIN_VERTICES vertices, 2^25
// by default) unless the caller chose one.
uint32_t merge_policy(int32_t sx, int32_t sy, uint32_t flags);

// Out-of-core mode: spill_piece stores piece i (when live) and frees it;
// restore_piece makes it live again and returns it.  Both are no-ops for a
// piece in its target state; tasks on different pieces may run together.
void spill_piece(mscx_grid *G, int i);
mscx_complex *restore_piece(mscx_grid *G, int i);

}  // namespace mscx
