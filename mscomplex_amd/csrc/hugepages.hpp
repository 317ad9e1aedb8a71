// Transparent huge pages for large host buffers.
#pragma once

#include <sys/mman.h>

#include <cstddef>
#include <cstdint>

namespace mscx {

// Advises the whole 2 MiB pages inside [p, p + bytes) onto transparent huge
// pages where the system allows it (madvise mode): a buffer advised before
// its first touch then takes one fault per 2 MiB instead of per 4 KiB, and
// random accesses over it miss the TLB less.  A hint; failure is ignored.
inline void advise_huge(const void *p, size_t bytes) {
  constexpr uintptr_t H = (uintptr_t)2 << 20;
  if (!p || bytes < 2 * H) return;
  const uintptr_t b = ((uintptr_t)p + H - 1) & ~(H - 1), e = ((uintptr_t)p + bytes) & ~(H - 1);
  if (e > b) (void)madvise((void *)b, e - b, MADV_HUGEPAGE);
}

}  // namespace mscx
