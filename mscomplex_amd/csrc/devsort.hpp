// Device primitives shared by the merge tree and the manifold pass
// (devsort.hip): an exclusive scan of 32-bit words and a stable LSD radix
// sort of (key, value) word pairs.  Both run on the caller's stream with
// scratch the caller provides (no allocation inside).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace mscx {

// Words of scratch dev_scan needs for n items.
uint64_t scan_temp_words(uint64_t n);
// Exclusive scan of n words (in and out may be the same array); the total
// lands in *d_total when d_total is not null.
void dev_scan(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *d_total, uint32_t *tmp, hipStream_t s);

// Words of scratch radix_sort_pairs needs for n pairs (besides the two
// alternate key/value arrays).
uint64_t sort_temp_words(uint64_t n);
// Stable sort of n pairs by their keys, of which only the low `bits` bits
// may be set: pairs of equal keys keep their input order.  keys/vals hold
// the input; the result is in (keys, vals) or in (keys2, vals2) -- the
// return value says which (false: keys/vals).  8-bit digits, one pass per
// digit: block histograms, a scan over (digit, block), and a scatter that
// ranks equal digits within a block by wave ballots in input order.
bool radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys2, uint32_t *vals2, uint64_t n, int bits,
                      uint32_t *tmp, hipStream_t s);

}  // namespace mscx
