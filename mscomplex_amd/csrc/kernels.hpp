// Host-side declarations of the kernel launchers (kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "tile_desc.hpp"

namespace mscx {

enum BlockKind { KB_GRADIENT = 0, KB_SCAN = 1, KB_ROWSEG = 2, KB_CRIT = 3, KB_COUNT = 4 };

// prefix[0..ntiles] = first block of each tile for a launch of `kind`;
// returns the total block count.
int plan_blocks(int kind, const TileDesc *tiles, int ntiles, int *prefix);

void launch_grad_trace(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s);
void launch_resolve_exits(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, uint32_t *d_err,
                          hipStream_t s);
void launch_scan(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s);
void launch_compact(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s);
void launch_incidence(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s);

// Manifold labels for ndisc leaves; prefix has ndisc+1 entries of 256-item blocks.
int plan_disc_blocks(const DiscDesc *d, int n, int *prefix);
void launch_disc_labels(const DiscDesc *d_desc, const int *d_prefix, int n, int nblocks, hipStream_t s);

}  // namespace mscx
