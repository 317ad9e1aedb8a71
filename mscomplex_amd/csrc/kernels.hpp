// Host-side declarations of the kernel launchers (kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "tile_desc.hpp"

namespace mscx {

enum BlockKind { KB_GRADIENT = 0, KB_SCAN = 1, KB_RING = 2, KB_COMPACT = 3, KB_CRIT = 4, KB_COUNT = 5 };

// prefix[0..ntiles] = first block of each tile for a launch of `kind`;
// returns the total block count.
int plan_blocks(int kind, const TileDesc *tiles, int ntiles, int *prefix);

// Error flags set by the kernels in *d_err: 2 = a block-local trace did not
// end, 4 = a ring chain did not end, 8 = a path left the ext window (all
// mean an invalid gradient).
void launch_grad_local(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, uint32_t *d_err,
                       hipStream_t s);
void launch_scan(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s);
void launch_ring_resolve(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, uint32_t *d_err,
                         hipStream_t s);
void launch_owner(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s);
void launch_compact(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s);
void launch_incidence(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s);

// Disc tracing over the slots of n leaves (prefix: n+1 block offsets of
// 256 slots); write = false counts each slot's list, true writes it.
int plan_disc_trace_blocks(const DiscTraceDesc *d, int n, int *prefix);
void launch_disc_trace(const DiscTraceDesc *d_desc, const int *d_prefix, int n, int nblocks, bool write,
                       hipStream_t s);
// Manifold labels for ndisc leaves; prefix has ndisc+1 entries of 256-item blocks.
int plan_disc_blocks(const DiscDesc *d, int n, int *prefix);
void launch_disc_labels(const DiscDesc *d_desc, const int *d_prefix, int n, int nblocks, hipStream_t s);
// Sort input of the extremum discs' grouping over n leaves (prefix: n+1
// block offsets of 256 elements, vertices then quads of each rect).
int plan_disc_group_blocks(const DiscDesc *d, int n, int *prefix);
void launch_disc_group_keys(const DiscDesc *d_desc, const int *d_prefix, int n, int nblocks, uint32_t sentinel,
                            uint32_t *keys, uint32_t *vals, hipStream_t s);
// off[k] for k in [0, nkeys]: the first position of a key >= k in sorted[0, n).
void launch_disc_group_offsets(const uint32_t *sorted, uint64_t n, uint32_t nkeys, uint32_t *off, hipStream_t s);
// The stage chain of run_tiles up to the critical counts (upload of the
// descriptor block, gradient, scan, ring resolution, owner images, download
// of the counts) as a HIP graph built node by node, so that no stream is
// ever in capture mode.  d_prefix holds KB_COUNT prefix tables of ntiles+1.
// With ev != NULL the graph also records ev[0]/ev[1] around the gradient
// kernel, ev[2]/ev[3] around the scan, ev[4]/ev[5] around the ring
// resolution and ev[6]/ev[7] around the owner kernel (stage timing of the
// replayed graph).
hipError_t build_front_graph(hipGraph_t *out, void *up_dst, const void *up_src, size_t up_bytes, void *dl_dst,
                             const void *dl_src, size_t dl_bytes, const TileDesc *d_tiles, const int *d_prefix,
                             int ntiles, const int *nblocks, uint32_t *d_err, const hipEvent_t *ev);

}  // namespace mscx
