// mscx_dataset.hpp -- header-only C++ drop-in for the reference's GridDataset
// and GridMSComplex (nithins/mscomplex, dmsc) over the C ABI of libmscx.so.
//
// A dmsc maintainer replaces
//   dmsc/include/grid_dataset.h:35-245   (GridDataset)
//   dmsc/include/grid_mscomplex.h:48-84  (GridMSComplex)
// with these classes: the method names, the order in which
// GridDataManager calls them (grid_datamanager.cpp:121-210,397-451,596-662)
// and the host arrays they fill keep the reference's meaning and layout:
//   m_cell_flags  int8 flags (1 paired, 2 critical, 3 boundary-cancellable),
//   m_cell_pairs  pair of each cell ((-1,-1) when unpaired),
//   m_cell_own    owning extremum of each cell ((-1,-1) when unreached),
// each a 2-D array over the ext rect indexed [x][y] in C order (y fastest),
// as the reference's boost::multi_array (grid_dataset.cpp:1468-1483), and
//   m_critical_cells  the critical cells in raster order (y-major).
// work() runs the reference's CPU path (gradient, collation, owners) on the
// GPU; work_gpu() is the same call under the name the reference's OpenCL
// twin (work_ocl) suggests.  Failures rethrow the reference's exception
// types: std::logic_error for invariants (MSCX_ELOGIC), std::invalid_argument
// for bad arguments, std::runtime_error for device and file errors.
#ifndef MSCX_DATASET_HPP_INCLUDED
#define MSCX_DATASET_HPP_INCLUDED

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "mscx.h"

namespace mscx_dmsc {

inline void check(int rc) {
  if (rc == MSCX_OK) return;
  const std::string m = mscx_last_error();
  if (rc == MSCX_ELOGIC) throw std::logic_error(m);
  if (rc == MSCX_EINVAL) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}

// cellid_t of the reference (grid_mscomplex.h:31-45): two int16 cell
// coordinates, (-1,-1) when unset (rectangle_complex.h:69).
struct cellid_t {
  int16_t c[2] = {-1, -1};
  cellid_t() {}
  cellid_t(int x, int y) {
    c[0] = (int16_t)x;
    c[1] = (int16_t)y;
  }
  int16_t &operator[](int i) { return c[i]; }
  int16_t operator[](int i) const { return c[i]; }
  bool operator==(const cellid_t &o) const { return c[0] == o.c[0] && c[1] == o.c[1]; }
  bool operator!=(const cellid_t &o) const { return !(*this == o); }
  // the packed form of the C ABI (x | y << 16)
  static cellid_t unpack(uint32_t p) {
    return p == 0xFFFFFFFFu ? cellid_t() : cellid_t((int16_t)(p & 0xFFFF), (int16_t)(p >> 16));
  }
  uint32_t packed() const { return (uint32_t)(uint16_t)c[0] | ((uint32_t)(uint16_t)c[1] << 16); }
};

// rect_t of the reference (rectangle_complex.h:57-308): bottom-left and
// top-right cells, inclusive.
struct rect_t {
  cellid_t bl, tr;
  rect_t() {}
  rect_t(int x0, int y0, int x1, int y1) : bl(x0, y0), tr(x1, y1) {}
  rect_t(const cellid_t &a, const cellid_t &b) : bl(a), tr(b) {}
  int left() const { return bl[0]; }
  int bottom() const { return bl[1]; }
  int right() const { return tr[0]; }
  int top() const { return tr[1]; }
  bool contains(const cellid_t &c) const {
    return c[0] >= bl[0] && c[0] <= tr[0] && c[1] >= bl[1] && c[1] <= tr[1];
  }
  void as_array(int32_t out[4]) const {
    out[0] = bl[0];
    out[1] = bl[1];
    out[2] = tr[0];
    out[3] = tr[1];
  }
};

class GridDataset;

// GridMSComplex (grid_mscomplex.h:48-84): a host-resident complex owned by
// this object.
class GridMSComplex {
 public:
  rect_t m_rect, m_ext_rect;

  GridMSComplex(const rect_t &r, const rect_t &e) : m_rect(r), m_ext_rect(e) {}
  ~GridMSComplex() { reset(nullptr); }
  GridMSComplex(const GridMSComplex &) = delete;
  GridMSComplex &operator=(const GridMSComplex &) = delete;

  // grid_mscomplex.cpp:24-156; the caller owns the result
  static GridMSComplex *merge_up(const GridMSComplex &a, const GridMSComplex &b) {
    mscx_complex *h = nullptr;
    check(mscx_complex_merge_up(a.handle(), b.handle(), &h));
    GridMSComplex *m = new GridMSComplex(union_of(a.m_rect, b.m_rect), union_of(a.m_ext_rect, b.m_ext_rect));
    m->reset(h);
    return m;
  }
  // grid_mscomplex.cpp:158-322
  void merge_down(GridMSComplex &a, GridMSComplex &b) { check(mscx_complex_merge_down(handle(), a.handle(), b.handle())); }
  // grid_mscomplex.cpp:528-535
  void simplify_un_simplify(double threshold) { check(mscx_complex_simplify_un_simplify(handle(), threshold)); }
  // grid_mscomplex.cpp:573-590
  void write_discs(const std::string &prefix) const { check(mscx_complex_write_discs(handle(), prefix.c_str())); }

  size_t num_cps() const {
    uint32_t n = 0;
    check(mscx_complex_num_cps(handle(), &n));
    return n;
  }
  // the canonical dump words (format in mscx.h)
  std::vector<uint32_t> dump() const {
    size_t n = 0;
    check(mscx_complex_dump(handle(), nullptr, 0, &n));
    std::vector<uint32_t> w(n);
    check(mscx_complex_dump(handle(), w.data(), w.size(), &n));
    return w;
  }

  mscx_complex *handle() const {
    if (!h_) throw std::logic_error("complex has not been filled (writeout_connectivity / merge_up)");
    return h_;
  }
  void reset(mscx_complex *h) {
    if (h_) mscx_complex_destroy(h_);
    h_ = h;
  }

 private:
  mscx_complex *h_ = nullptr;
  static rect_t union_of(const rect_t &a, const rect_t &b) {
    return rect_t(std::min(a.left(), b.left()), std::min(a.bottom(), b.bottom()), std::max(a.right(), b.right()),
                  std::max(a.top(), b.top()));
  }
};

// GridDataset (grid_dataset.h:35-245): one rect + ext window on the GPU.
class GridDataset {
 public:
  rect_t m_rect, m_ext_rect;
  std::vector<int8_t> m_cell_flags;       // [x][y] over ext, y fastest
  std::vector<cellid_t> m_cell_pairs;     // [x][y] over ext, y fastest
  std::vector<cellid_t> m_cell_own;       // [x][y] over ext, y fastest
  std::vector<cellid_t> m_critical_cells; // raster order

  // grid_dataset.cpp:1429-1441
  GridDataset(const rect_t &r, const rect_t &e) : m_rect(r), m_ext_rect(e) {
    int32_t rr[4], ee[4];
    r.as_array(rr);
    e.as_array(ee);
    check(mscx_tile_create(rr, ee, &t_));
  }
  ~GridDataset() {
    if (t_) mscx_tile_destroy(t_);
  }
  GridDataset(const GridDataset &) = delete;
  GridDataset &operator=(const GridDataset &) = delete;

  // grid_dataset.h:236-238: the device is selected once per process
  static void init_opencl() { check(mscx_init(0)); }
  static void stop_opencl() { check(mscx_shutdown()); }

  // grid_dataset.cpp:1461-1484: the ext window of vertex values, x fastest
  // (Fortran order), (ext_w/2 + 1) x (ext_h/2 + 1) floats, copied to HBM
  void init(const float *data) {
    check(mscx_tile_upload(t_, data, (size_t)num_vert(0) * num_vert(1)));
    m_cell_flags.clear();
    m_cell_pairs.clear();
    m_cell_own.clear();
    m_critical_cells.clear();
  }
  // grid_dataset.cpp:1486-1509: the host arrays are released
  void clear() {
    std::vector<int8_t>().swap(m_cell_flags);
    std::vector<cellid_t>().swap(m_cell_pairs);
    std::vector<cellid_t>().swap(m_cell_own);
    std::vector<cellid_t>().swap(m_critical_cells);
  }

  // grid_dataset.cpp:1734-1741 (assignGradients, collateCriticalPoints,
  // assignCellOwnerExtrema) on the GPU, then the host arrays in the
  // reference's layout
  void work() { work_gpu(); }
  void work_gpu() {
    check(mscx_tile_work(t_));
    const size_t W = (size_t)cells(0), H = (size_t)cells(1), n = W * H;
    uint32_t K = 0;
    check(mscx_tile_num_critical(t_, &K));
    std::vector<int8_t> f(n);
    std::vector<uint32_t> p(n), o(n), c(K);
    check(mscx_tile_download(t_, f.data(), p.data(), o.data(), c.data()));
    m_cell_flags.resize(n);
    m_cell_pairs.resize(n);
    m_cell_own.resize(n);
    // the C ABI's images are x fastest; the reference's arrays are [x][y]
    for (size_t y = 0; y < H; ++y)
      for (size_t x = 0; x < W; ++x) {
        const size_t s = x + y * W, d = x * H + y;
        m_cell_flags[d] = f[s];
        m_cell_pairs[d] = cellid_t::unpack(p[s]);
        m_cell_own[d] = cellid_t::unpack(o[s]);
      }
    m_critical_cells.resize(K);
    for (uint32_t k = 0; k < K; ++k) m_critical_cells[k] = cellid_t::unpack(c[k]);
  }

  // grid_dataset.cpp:1858-1907: the complex of this tile into msc
  void writeout_connectivity(GridMSComplex *msc) {
    mscx_complex *h = nullptr;
    check(mscx_complex_from_tile(t_, &h));
    msc->reset(h);
  }
  // grid_dataset.cpp:1044-1231
  int postMergeFillDiscs(GridMSComplex *msc) {
    check(mscx_tile_fill_discs(t_, msc->handle()));
    return 0;
  }

  // element accessors in the reference's indexing (absolute cell coords)
  int8_t flag(const cellid_t &c) const { return m_cell_flags[index(c)]; }
  cellid_t pair(const cellid_t &c) const { return m_cell_pairs[index(c)]; }
  cellid_t owner(const cellid_t &c) const { return m_cell_own[index(c)]; }
  size_t index(const cellid_t &c) const {
    return (size_t)(c[0] - m_ext_rect.left()) * cells(1) + (size_t)(c[1] - m_ext_rect.bottom());
  }
  int cells(int axis) const { return (axis == 0 ? m_ext_rect.right() - m_ext_rect.left()
                                                : m_ext_rect.top() - m_ext_rect.bottom()) + 1; }
  int num_vert(int axis) const { return cells(axis) / 2 + 1; }

  mscx_tile *handle() const { return t_; }

 private:
  mscx_tile *t_ = nullptr;
};

}  // namespace mscx_dmsc

#endif  // MSCX_DATASET_HPP_INCLUDED
