// Device-resident Morse-Smale complexes and the quadtree merge tree on the
// GPU (dmerge.hip).
//
// A complex lives in HBM as flat per-node arrays (cell, value, pair, flags),
// two arc lists per node (ascending, descending) in one pool of 32-bit node
// indices -- list k of node i is arcs[off[2i+k], off[2i+k] + cnt[2i+k]),
// sorted, duplicates kept (the reference's std::multiset<uint>) -- and an
// open-addressing cell -> index map of 64-bit slots (key << 32 | index).
//
// merge_up and merge_down (grid_mscomplex.cpp:24-322) run as kernels:
// union index assignment by scans, node and arc copies through index maps
// (gathers), per-list sorts, replica ordering by first reference.  The
// cancellations along the shared line (cancelPairs / uncancel_pairs,
// discreteMorseAlgorithm.h:348-464,1006-1089) form one dependency chain --
// each pair's rewiring changes the lists the next pair reads, and the lists
// a cancelled pair keeps are observable -- so the line's sub-complex (its
// nodes, their pairs and neighbours) is gathered on the device, cancelled on
// a host thread with the host algebra, and the changed nodes are written
// back into the pool.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "complex.hpp"

struct mscx_tile;

namespace mscx {

struct DComplex {
  Rect rect{}, ext{};
  uint32_t n = 0, ncap = 0;  // nodes, node capacity
  uint64_t aused = 0, acap = 0;  // arc pool words used / allocated
  uint32_t mmask = 0;  // map slots - 1
  uint32_t *cell = nullptr, *pair = nullptr;
  float *fn = nullptr;
  uint8_t *flags = nullptr;
  uint32_t *off = nullptr, *cnt = nullptr;  // 2 per node
  uint32_t *arcs = nullptr;
  unsigned long long *map = nullptr;
  void *node_mem = nullptr, *arc_mem = nullptr, *map_mem = nullptr;
  DComplex() = default;
  DComplex(const DComplex &) = delete;
  DComplex &operator=(const DComplex &) = delete;
  ~DComplex();
  size_t device_bytes() const;
};

// Raised (as a return value) when a merge meets a case whose reference
// semantics the device kernels do not follow: a pair lookup that would
// insert a zero entry into the reference's cell map, an arc list longer
// than one workgroup sorts, a non-mutual pairing in merge_down.  The caller
// then runs the whole grid's tree on the host.
struct DFallback {
  std::string why;
  explicit DFallback(std::string w = std::string()) : why(std::move(w)) {}
};
extern std::atomic<uint64_t> g_dfallbacks;

// Leaf complex (writeout_connectivity) from a worked tile's critical block.
void dcomplex_from_tile(const mscx_tile *t, const Rect &rect, const Rect &ext, DComplex &out, hipStream_t s);
// merge_up(a, b) into out; throws DFallback or the reference's logic_error.
void dmerge_up(const DComplex &a, const DComplex &b, DComplex &out, hipStream_t s);
// merge_down(P, a, b); P's line pairs are uncancelled in place.
void dmerge_down(DComplex &P, DComplex &a, DComplex &b, hipStream_t s);
// Device copy (same capacities) on stream s.
void dcomplex_copy(const DComplex &src, DComplex &dst, hipStream_t s);
// Host copies.
void dcomplex_download(const DComplex &d, Complex &out, hipStream_t s);
void dcomplex_upload(const Complex &c, DComplex &out, hipStream_t s);

// A non-blocking stream (default priority) from a pool, for the duration of
// one task of a parallel loop.
struct StreamLease {
  hipStream_t s = nullptr;
  StreamLease();
  ~StreamLease();
  StreamLease(const StreamLease &) = delete;
  StreamLease &operator=(const StreamLease &) = delete;
};

}  // namespace mscx
