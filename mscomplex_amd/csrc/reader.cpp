// Raw-grid reader of the file-driven run: the counterpart of
// GridDataManager::readDataAndInit / computeSubdomainMsgraphs
// (grid_datamanager.cpp:493-594).
//
// The reference reads the rows of num_parallel pieces into one of two host
// buffers, computes those pieces while the next buffer is read, and seeks
// back three vertex rows between buffers because consecutive pieces' ext
// rects share three rows (the split row and one halo row either side).
// Here the whole grid (at most 16384^2 floats, 1 GiB) becomes resident in
// HBM: the file is read front to back once, in row slabs, into two pinned
// host buffers, and each slab's copy to the device runs on the work stream
// while the next slab is read.  The pieces then bind to the resident grid
// at their ext offsets, so every row is read once and nothing is rewound.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <cstring>

#include "pipeline.hpp"

namespace mscx {

GridFile::~GridFile() {
  if (fd >= 0) ::close(fd);
}

void open_grid_file(const char *path, int32_t sx, int32_t sy, GridFile &f) {
  f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
  MSCX_REQUIRE(f.fd >= 0, MSCX_EIO, std::string("unable to read stream: ") + path + ": " + strerror(errno));
  struct stat st;
  MSCX_REQUIRE(fstat(f.fd, &st) == 0, MSCX_EIO, std::string("unable to read stream: ") + path + ": " + strerror(errno));
  f.bytes = (size_t)sx * (size_t)sy * 4;
  MSCX_REQUIRE(S_ISREG(st.st_mode) || S_ISBLK(st.st_mode) || S_ISCHR(st.st_mode) || S_ISFIFO(st.st_mode), MSCX_EIO,
               std::string("not a data file: ") + path);
  MSCX_REQUIRE(!S_ISREG(st.st_mode) || (size_t)st.st_size >= f.bytes, MSCX_EIO,
               std::string("file holds fewer than size_x * size_y floats: ") + path);
}

// Reads exactly n bytes at the file's current position.
static void read_full(int fd, char *p, size_t n) {
  while (n) {
    const ssize_t r = ::read(fd, p, std::min<size_t>(n, (size_t)1 << 30));
    if (r < 0 && errno == EINTR) continue;
    MSCX_REQUIRE(r > 0, MSCX_EIO, r == 0 ? "file ended before size_x * size_y floats" : strerror(errno));
    p += r;
    n -= (size_t)r;
  }
}

static size_t slab_bytes(size_t row_bytes) {
  size_t want = (size_t)8 << 20;
  if (const char *e = getenv("MSCX_READ_SLAB_BYTES")) {
    const long long v = atoll(e);
    if (v > 0) want = (size_t)v;
  }
  return std::max<size_t>(1, want / row_bytes) * row_bytes;
}

void stream_grid_file(GridFile &f, int32_t sx, float *dev) {
  Runtime &rt = runtime();
  const size_t row = (size_t)sx * 4;
  const size_t slab = std::min(slab_bytes(row), f.bytes);
  struct Staging {
    hipStream_t stream;
    char *buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    ~Staging() {
      // copies still in flight read the buffers: wait before freeing them
      (void)hipStreamSynchronize(stream);
      for (int k = 0; k < 2; ++k) {
        if (done[k]) (void)hipEventDestroy(done[k]);
        if (buf[k]) (void)hipHostFree(buf[k]);
      }
    }
  } st{rt.stream};
  for (int k = 0; k < 2; ++k) {
    MSCX_HIP(hipHostMalloc((void **)&st.buf[k], slab, hipHostMallocDefault));
    MSCX_HIP(hipEventCreateWithFlags(&st.done[k], hipEventDisableTiming));
  }
  bool issued[2] = {false, false};
  int k = 0;
  for (size_t off = 0; off < f.bytes; off += slab, k ^= 1) {
    const size_t n = std::min(slab, f.bytes - off);
    if (issued[k]) MSCX_HIP(hipEventSynchronize(st.done[k]));  // the copy two slabs back
    read_full(f.fd, st.buf[k], n);
    // NaN has no place in the vertex order; checked as bits so the loop vectorises
    const uint32_t *u = (const uint32_t *)st.buf[k];
    uint32_t nan = 0;
    for (size_t i = 0; i < n / 4; ++i) nan |= (uint32_t)((u[i] & 0x7FFFFFFFu) > 0x7F800000u);
    MSCX_REQUIRE(!nan, MSCX_EINVAL, "input contains NaN");
    MSCX_HIP(hipMemcpyAsync((char *)dev + off, st.buf[k], n, hipMemcpyHostToDevice, rt.stream));
    MSCX_HIP(hipEventRecord(st.done[k], rt.stream));
    issued[k] = true;
  }
  MSCX_HIP(hipStreamSynchronize(rt.stream));
}

}  // namespace mscx
