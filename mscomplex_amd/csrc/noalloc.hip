// mscx_tile_work_noalloc: GridDataset::work() on one tile with no device
// allocation.  The tile's intermediates (gradient codes, segment counts and
// offsets, critical bits, owner image, local terminals, ring words) and the
// launch descriptor live in `scratch`, a file-scope device array of
// NOALLOC_SCRATCH_BYTES; the caller passes the input window and the output
// arrays of the critical cells.  Every call on a device shares that
// device's array (the library's launch lock serialises them).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "kernels.hpp"
#include "runtime.hpp"
#include "tile.hpp"

namespace mscx {

// 32 MiB: the intermediates of a tile of up to about 1026 x 1026 vertices
// (a 4096^2 grid's 4x4 leaves with their halos take 29.4 MB).
constexpr size_t NOALLOC_SCRATCH_BYTES = 32u << 20;
constexpr size_t NOALLOC_HEAD = 4096;  // descriptor, block prefixes, K, error flag
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Per-device address of `scratch`, looked up once per device with that
// device current (the array's host-side name is not a device address).
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static unsigned char *scratch_ptr(int dev) {
  MSCX_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, MSCX_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    MSCX_REQUIRE(e == hipSuccess && p, MSCX_EDEVICE,
                 std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  return (unsigned char *)g_scratch_ptr[dev];
}

struct NoallocHead {
  TileDesc desc;
  int prefix[KB_COUNT][2];
  uint32_t k, err;
};
static_assert(sizeof(NoallocHead) <= NOALLOC_HEAD, "descriptor head too large");

// Writes the launch descriptor and block prefixes into the scratch head;
// clears K and the error flag when `clear` is set.  A fill kernel rather
// than a memset or copy, so that the call needs no host staging.
__global__ void k_noalloc_head(NoallocHead h, NoallocHead *dst, int clear) {
  if (threadIdx.x != 0) return;
  dst->desc = h.desc;
  for (int k = 0; k < KB_COUNT; ++k) {
    dst->prefix[k][0] = h.prefix[k][0];
    dst->prefix[k][1] = h.prefix[k][1];
  }
  if (clear) {
    dst->k = 0;
    dst->err = 0;
  }
}

}  // namespace mscx

using namespace mscx;

extern "C" int mscx_tile_work_noalloc(const int32_t rect[4], const int32_t ext[4], const float *dev_verts,
                                      int64_t row_pitch, uint32_t kcap, uint32_t *dev_crit, float *dev_fn,
                                      uint32_t *dev_conn4, uint32_t *dev_pair, uint32_t *K) {
  return guarded([&] {
    MSCX_REQUIRE(rect && ext && dev_verts && K, MSCX_EINVAL, "null argument");
    check_rects(rect, ext);
    ensure_init();
    Runtime &rt = runtime();
    NoallocHead h{};
    size_t o[7];
    const size_t body = fixed_layout(rect, ext, h.desc, o);
    MSCX_REQUIRE(NOALLOC_HEAD + body <= NOALLOC_SCRATCH_BYTES, MSCX_EINVAL,
                 "tile too large for the no-allocation scratch (" + std::to_string(NOALLOC_HEAD + body) +
                     " bytes needed, " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                     " available: at most about 1026 x 1026 vertices)");
    MSCX_REQUIRE(row_pitch >= h.desc.g.vx, MSCX_EINVAL, "row pitch smaller than the window width");
    MSCX_REQUIRE(kcap == 0 || (dev_crit && dev_fn && dev_conn4 && dev_pair), MSCX_EINVAL,
                 "null output array");
    std::lock_guard<std::mutex> lk(rt.mu);
    hipStream_t s = rt.stream;
    unsigned char *base = scratch_ptr(rt.device);
    NoallocHead *dh = (NoallocHead *)base;
    unsigned char *b = base + NOALLOC_HEAD;
    TileDesc &d = h.desc;
    d.verts = dev_verts;
    d.vpitch = row_pitch;
    d.codes = (uint8_t *)(b + o[0]);
    d.seg_offs = (uint32_t *)(b + o[1]);
    d.scan_part = (uint32_t *)(b + o[2]);
    d.critbits = (uint64_t *)(b + o[3]);
    d.own = (uint32_t *)(b + o[4]);
    d.loc = (uint32_t *)(b + o[5]);
    d.ring = (uint32_t *)(b + o[6]);
    d.ringidx = d.ring + ring_words_max(d.g, d.nseg);
    d.nkrit = &dh->k;
    d.kcap = 0;
    int nb[KB_COUNT] = {};
    for (int k = 0; k < KB_CRIT; ++k) nb[k] = plan_blocks(k, &d, 1, h.prefix[k]);
    const TileDesc *dd = &dh->desc;
    const int *dp = &dh->prefix[0][0];
    uint32_t *derr = &dh->err;

    hipLaunchKernelGGL(k_noalloc_head, dim3(1), dim3(64), 0, s, h, dh, 1);
    launch_grad_local(dd, dp + 2 * KB_GRADIENT, 1, nb[KB_GRADIENT], derr, s);
    launch_scan(dd, dp + 2 * KB_SCAN, 1, nb[KB_SCAN], s);
    launch_ring_resolve(dd, dp + 2 * KB_RING, 1, nb[KB_RING], derr, s);
    launch_owner(dd, dp + 2 * KB_GRADIENT, 1, nb[KB_GRADIENT], s);
    MSCX_HIP(hipGetLastError());
    uint32_t ke[2] = {0, 0};
    MSCX_HIP(hipMemcpyAsync(ke, &dh->k, sizeof(ke), hipMemcpyDeviceToHost, s));
    MSCX_HIP(hipStreamSynchronize(s));
    MSCX_REQUIRE(ke[1] == 0, MSCX_ELOGIC, "manifold trace did not resolve (invalid gradient)");
    *K = ke[0];
    MSCX_REQUIRE(ke[0] <= kcap, MSCX_EINVAL,
                 "critical-cell capacity too small: " + std::to_string(ke[0]) + " needed (written to *K)");

    d.crit = dev_crit;
    d.crit_fn = dev_fn;
    d.crit_conn = dev_conn4;
    d.crit_pair = dev_pair;
    d.kcap = ke[0];
    nb[KB_CRIT] = plan_blocks(KB_CRIT, &d, 1, h.prefix[KB_CRIT]);
    hipLaunchKernelGGL(k_noalloc_head, dim3(1), dim3(64), 0, s, h, dh, 0);
    launch_compact(dd, dp + 2 * KB_COMPACT, 1, nb[KB_COMPACT], s);
    launch_incidence(dd, dp + 2 * KB_CRIT, 1, nb[KB_CRIT], s);
    MSCX_HIP(hipGetLastError());
    MSCX_HIP(hipStreamSynchronize(s));
  });
}
