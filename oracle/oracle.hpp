// oracle/oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference's CPU path (dmsc GridDataset::work(),
// writeout_connectivity, postMergeFillDiscs, MSComplex cancellation algebra,
// GridMSComplex merge/simplify, GridDataManager pipeline order).  It exists
// to check the product (the HIP path under mscomplex_amd/) and to time the
// CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may link or call it; product code never does.
//
// Each routine cites the reference file:line it follows.  The containers
// (std::map, std::multiset, std::priority_queue, std::queue) are the same
// libstdc++ containers the reference uses, because their iteration and heap
// orders are observable in the output (multiset arc multiplicities,
// equal-persistence cancellation order).
#pragma once

#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

namespace orc {

extern int g_plant;  // planted mistake for the pin tests (0: none)
extern int g_stop_after;  // last pipeline phase run (4: all)


typedef int16_t coord_t;
typedef float fn_t;
typedef unsigned int uint;

// cellid_t / point_def: rectangle_complex.h:57-106 (two_tuple_t of int16).
// Default value is (-1,-1) (rectangle_complex.h:69).  Ordering is
// lexicographic on (x, y), the order of a boost::array-derived tuple.
struct cellid_t {
  coord_t v[2];
  cellid_t() { v[0] = -1; v[1] = -1; }
  cellid_t(int x, int y) { v[0] = (coord_t)x; v[1] = (coord_t)y; }
  coord_t &operator[](int i) { return v[i]; }
  coord_t operator[](int i) const { return v[i]; }
  bool operator<(const cellid_t &o) const {
    return v[0] < o.v[0] || (v[0] == o.v[0] && v[1] < o.v[1]);
  }
  bool operator>(const cellid_t &o) const { return o < *this; }
  bool operator==(const cellid_t &o) const { return v[0] == o.v[0] && v[1] == o.v[1]; }
  bool operator!=(const cellid_t &o) const { return !(*this == o); }
  cellid_t half() const { return cellid_t(v[0] / 2, v[1] / 2); }
  uint32_t packed() const { return (uint16_t)v[0] | ((uint32_t)(uint16_t)v[1] << 16); }
};

// rectangle_def: rectangle_complex.h:108-308 (only the predicates used on
// the grid path).
struct rect_t {
  cellid_t bl, tr;
  rect_t() {}
  rect_t(const cellid_t &p1, const cellid_t &p2) {
    bl = cellid_t(std::min(p1[0], p2[0]), std::min(p1[1], p2[1]));
    tr = cellid_t(std::max(p1[0], p2[0]), std::max(p1[1], p2[1]));
  }
  rect_t(int x0, int y0, int x1, int y1) : rect_t(cellid_t(x0, y0), cellid_t(x1, y1)) {}
  bool isInInterior(const cellid_t &p) const {  // :158-168
    return p[0] > bl[0] && p[1] > bl[1] && p[0] < tr[0] && p[1] < tr[1];
  }
  bool contains(const cellid_t &p) const {  // :170-179
    return p[0] >= bl[0] && p[1] >= bl[1] && p[0] <= tr[0] && p[1] <= tr[1];
  }
  bool isOnBoundry(const cellid_t &p) const { return contains(p) && !isInInterior(p); }  // :181-184
  bool intersects(const rect_t &r) const {  // :197-206
    return !(tr[0] < r.bl[0] || r.tr[0] < bl[0] || tr[1] < r.bl[1] || r.tr[1] < bl[1]);
  }
  bool intersection(const rect_t &r, rect_t &i) const {  // :208-220
    if (!intersects(r)) return false;
    int l = std::max(r.left(), left()), rr = std::min(r.right(), right());
    int b = std::max(r.bottom(), bottom()), t = std::min(r.top(), top());
    i = rect_t(l, b, rr, t);
    return true;
  }
  int left() const { return bl[0]; }
  int right() const { return tr[0]; }
  int bottom() const { return bl[1]; }
  int top() const { return tr[1]; }
  cellid_t size() const { return cellid_t(tr[0] - bl[0], tr[1] - bl[1]); }
};

// MSComplex<cellid_t>::critical_point -- discreteMorseDS.h:90-126
struct critpt_t {
  typedef std::multiset<uint> conn_t;
  typedef std::vector<cellid_t> disc_t;
  cellid_t cellid;
  uint pair_idx = (uint)-1;
  bool isCancelled = false;
  bool isOnStrangulationPath = false;
  bool isBoundryCancelable = false;
  disc_t asc_disc, des_disc;
  conn_t asc, des;
};

// MSComplex + GridMSComplex -- discreteMorseDS.h:78-143, grid_mscomplex.h:48-84
struct GridMSComplex {
  typedef std::pair<uint, uint> crit_idx_pair_t;
  typedef std::vector<crit_idx_pair_t> crit_idx_pair_list_t;

  std::vector<critpt_t *> m_cps;
  std::map<cellid_t, uint> m_id_cp_map;
  rect_t m_rect, m_ext_rect;
  std::vector<fn_t> m_cp_fns;

  GridMSComplex() {}
  GridMSComplex(rect_t r, rect_t e) : m_rect(r), m_ext_rect(e) {}
  ~GridMSComplex() { clear(); }
  void clear();

  static GridMSComplex *merge_up(const GridMSComplex &msc1, const GridMSComplex &msc2);
  void merge_down(GridMSComplex &msc1, GridMSComplex &msc2);
  void simplify(crit_idx_pair_list_t &canc_pairs_list, double thresh);
  void un_simplify(const crit_idx_pair_list_t &canc_pairs_list);
  void simplify_un_simplify(double thresh);

  // Flat dump used for comparisons (format shared with the product):
  //   [ncp, rect bl.x, bl.y, tr.x, tr.y, ext bl.x, bl.y, tr.x, tr.y]
  //   per cp: cellid, fn bits, pair_idx, flags(b0 cancelled, b1 strangulation,
  //           b2 boundary-cancelable), n_asc, n_des, n_asc_disc, n_des_disc,
  //           asc..., des..., asc_disc..., des_disc...
  void dump(std::vector<uint32_t> &out) const;
};

void cancelPairs(GridMSComplex *msc, uint cp0_ind, uint cp1_ind, std::vector<uint> *new_edges);
void uncancel_pairs(GridMSComplex *msc, uint cp1_idx, uint cp2_idx);

// GridDataset CPU engine -- grid_dataset.h:35-240, grid_dataset.cpp:1029-1951
class GridDataset {
 public:
  enum { CELLFLAG_UNKNOWN = 0, CELLFLAG_PAIRED = 1, CELLFLAG_CRITCAL = 2 };

  rect_t m_rect, m_ext_rect;
  const fn_t *m_verts = nullptr;  // (ext.x/2+1) x (ext.y/2+1), x fastest
  int m_vx = 0, m_vy = 0;         // vertex extents
  int m_cx = 0, m_cy = 0;         // cell extents
  std::vector<int8_t> m_cell_flags;
  std::vector<cellid_t> m_cell_pairs, m_cell_own;
  std::vector<cellid_t> m_critical_cells;

  GridDataset(const rect_t &r, const rect_t &e) : m_rect(r), m_ext_rect(e) {}

  void init(const fn_t *data);
  void clear();

  void work();
  void assignGradients();
  void collateCriticalPoints();
  void assignCellOwnerExtrema();
  void writeout_connectivity(GridMSComplex *msgraph);
  int postMergeFillDiscs(GridMSComplex *msgraph);

  // dataset interface
  size_t cidx(const cellid_t &c) const {
    return (size_t)(c[0] - m_ext_rect.bl[0]) * m_cy + (size_t)(c[1] - m_ext_rect.bl[1]);
  }
  fn_t vert(const cellid_t &c) const {
    int i = (c[0] >> 1) - (m_ext_rect.bl[0] >> 1), j = (c[1] >> 1) - (m_ext_rect.bl[1] >> 1);
    return m_verts[(size_t)i + (size_t)j * m_vx];
  }
  bool ptLt(cellid_t c1, cellid_t c2) const;
  bool compareCells(cellid_t c1, cellid_t c2) const;
  cellid_t getCellPairId(cellid_t c) const { return m_cell_pairs[cidx(c)]; }
  uint getCellPoints(cellid_t c, cellid_t *p) const;
  uint getCellFacets(cellid_t c, cellid_t *f) const;
  uint getCellCofacets(cellid_t c, cellid_t *cf) const;
  uint getCellIncCells(cellid_t c, cellid_t *inc) const;
  bool isCellMarked(cellid_t c) const { return m_cell_flags[cidx(c)] != CELLFLAG_UNKNOWN; }
  bool isCellCritical(cellid_t c) const { return m_cell_flags[cidx(c)] & CELLFLAG_CRITCAL; }
  bool isCellPaired(cellid_t c) const { return m_cell_flags[cidx(c)] & CELLFLAG_PAIRED; }
  void pairCells(cellid_t c, cellid_t p);
  void markCellCritical(cellid_t c) { m_cell_flags[cidx(c)] |= CELLFLAG_CRITCAL; }
  static uint getCellDim(cellid_t c) { return (c[0] & 1) + (c[1] & 1); }
  bool isTrueBoundryCell(cellid_t c) const { return m_ext_rect.isOnBoundry(c); }
  bool isCellExterior(cellid_t c) const { return !m_rect.contains(c) && m_ext_rect.contains(c); }
  fn_t get_cell_fn(cellid_t c) const;
};

// GridDataManager pipeline order -- grid_datamanager.cpp:40-747 (compute
// path only; rendering, UI and the out-of-core spill are not restated).
struct Piece {
  GridDataset *dataset = nullptr;
  GridMSComplex *msgraph = nullptr;
  uint level = 0;
  std::vector<fn_t> data;  // the piece's ext window of the input, x fastest
};

struct PipelineSnapshots {
  // Dumps of every piece's complex after each phase (index = piece number).
  std::vector<std::vector<uint32_t>> after_compute, after_merge_up, after_simplify,
      after_merge_down, after_discs;
  double t_compute = 0, t_merge_up = 0, t_simplify = 0, t_merge_down = 0, t_manifolds = 0;
};

class DataManager {
 public:
  // tiling: levels = slab levels along y as in createPieces_quadtree
  // (grid_datamanager.cpp:53-89), or, when tiles_lx/tiles_ly are given, a
  // 2-D tiling of 2^lx x 2^ly tiles built with the same split/halo rule.
  DataManager(const fn_t *data, int size_x, int size_y, int levels, int lx, int ly,
              double simp_tresh, bool keep_snapshots);
  ~DataManager();
  // nthreads > 1: independent pieces of one phase (leaf tiles, merges of
  // one tree level) run on that many threads, as with the reference's
  // --num-parallel; results do not depend on it.
  void run(PipelineSnapshots *snaps, int nthreads = 1);

  const fn_t *m_data;
  int m_size_x, m_size_y;
  int m_num_levels;
  int m_lx, m_ly;
  bool m_two_d;
  double m_simp_tresh;
  bool m_keep;
  std::vector<Piece *> m_pieces;

 private:
  void createPieces(rect_t r, rect_t e, int level, int lx, int ly);
  void initPiece(Piece *dp);
  void snapshot(std::vector<std::vector<uint32_t>> &dst);
};

}  // namespace orc
