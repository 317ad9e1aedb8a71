// Host-resident Morse-Smale complex (the product's GridMSComplex).
//
// Flat arrays per critical point instead of the reference's heap-allocated
// critical_point objects; arcs are sorted small multisets (iteration order
// and multiplicities identical to std::multiset<uint>); cell id -> index is
// an open-addressing hash map.  The algorithms follow
// dmsc/src/grid_mscomplex.cpp and dmsc/include/discreteMorseAlgorithm.h
// step for step, because tie orders, arc multiplicities and the
// strangulation flags they set are observable in the output.
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "hugepages.hpp"

struct mscx_tile;

namespace mscx {

// Sorted multiset of uint32 with inline storage for small sizes.
class ConnSet {
 public:
  ConnSet() {}
  ConnSet(const ConnSet &o) { assign(o); }
  ConnSet &operator=(const ConnSet &o) {
    if (this != &o) assign(o);
    return *this;
  }
  ConnSet(ConnSet &&o) noexcept { steal(o); }
  ConnSet &operator=(ConnSet &&o) noexcept {
    if (this != &o) {
      release();
      steal(o);
    }
    return *this;
  }
  ~ConnSet() { release(); }

  uint32_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  const uint32_t *begin() const { return data(); }
  const uint32_t *end() const { return data() + n_; }
  uint32_t operator[](uint32_t i) const { return data()[i]; }

  void clear() { n_ = 0; }
  // multiset::insert: after any equal elements
  void insert(uint32_t v) {
    reserve(n_ + 1);
    uint32_t *d = data();
    uint32_t pos = upper(v);
    memmove(d + pos + 1, d + pos, (n_ - pos) * 4);
    d[pos] = v;
    ++n_;
  }
  // multiset::erase(key): removes every element equal to v
  void erase_all(uint32_t v) {
    uint32_t *d = data();
    uint32_t lo = lower(v), hi = upper(v);
    if (lo == hi) return;
    memmove(d + lo, d + hi, (n_ - hi) * 4);
    n_ -= hi - lo;
  }
  uint32_t count(uint32_t v) const { return upper(v) - lower(v); }
  bool contains(uint32_t v) const {
    uint32_t l = lower(v);
    return l < n_ && data()[l] == v;
  }
  // append a value known to be >= every element (bulk build)
  void push_sorted(uint32_t v) {
    reserve(n_ + 1);
    data()[n_++] = v;
  }

 private:
  static constexpr uint32_t INL = 6;
  uint32_t n_ = 0, cap_ = INL;
  union {
    uint32_t inl_[INL];
    uint32_t *heap_;
  };
  uint32_t *data() { return cap_ > INL ? heap_ : inl_; }
  const uint32_t *data() const { return cap_ > INL ? heap_ : inl_; }
  uint32_t lower(uint32_t v) const {
    const uint32_t *d = data();
    uint32_t lo = 0, hi = n_;
    while (lo < hi) {
      uint32_t m = (lo + hi) >> 1;
      if (d[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
  }
  uint32_t upper(uint32_t v) const {
    const uint32_t *d = data();
    uint32_t lo = 0, hi = n_;
    while (lo < hi) {
      uint32_t m = (lo + hi) >> 1;
      if (d[m] <= v) lo = m + 1; else hi = m;
    }
    return lo;
  }
  void reserve(uint32_t want) {
    if (want <= cap_) return;
    uint32_t nc = cap_ * 2;
    if (nc < want) nc = want;
    uint32_t *nd = new uint32_t[nc];
    memcpy(nd, data(), n_ * 4);
    if (cap_ > INL) delete[] heap_;
    heap_ = nd;
    cap_ = nc;
  }
  void release() {
    if (cap_ > INL) delete[] heap_;
    cap_ = INL;
    n_ = 0;
  }
  void assign(const ConnSet &o) {
    n_ = 0;
    reserve(o.n_);
    memcpy(data(), o.data(), o.n_ * 4);
    n_ = o.n_;
  }
  void steal(ConnSet &o) {
    n_ = o.n_;
    cap_ = o.cap_;
    if (o.cap_ > INL)
      heap_ = o.heap_;
    else
      memcpy(inl_, o.inl_, sizeof(inl_));
    o.cap_ = INL;
    o.n_ = 0;
  }
};

// packed cell id (x | y << 16) -> critical-point index.  Open addressing
// with linear probing; a slot holds its key and value side by side, so a
// probe touches one cache line.
class CellMap {
 public:
  void reserve(size_t n);
  // returns the index, or UINT32_MAX when absent
  uint32_t find(uint32_t key) const;
  bool contains(uint32_t key) const { return find(key) != UINT32_MAX; }
  // map::insert semantics: keeps an existing value
  void insert(uint32_t key, uint32_t val) { find_or_insert(key, val); }
  // map::operator[] = val semantics: overwrites
  void set(uint32_t key, uint32_t val);
  // map::operator[] read semantics: inserts 0 when absent
  uint32_t get_or_zero(uint32_t key);
  // one probe: the existing value, or UINT32_MAX after inserting val
  uint32_t find_or_insert(uint32_t key, uint32_t val);
  // starts loading key's home slot (loops over many keys call it a few
  // keys ahead, so that the probes' cache misses overlap)
  void prefetch(uint32_t key) const {
    if (!slots_.empty()) __builtin_prefetch(&slots_[hash(key) & mask_]);
  }
  static uint32_t hash(uint32_t k) {
    k ^= k >> 16;
    k *= 0x7FEB352Du;
    k ^= k >> 15;
    k *= 0x846CA68Bu;
    k ^= k >> 16;
    return k;
  }
  size_t size() const { return n_; }
  // A table filled elsewhere with the same layout (home slot hash(key) &
  // (slots - 1), linear probing, {key, value} pairs, key UINT32_MAX empty):
  // returns room for `slots` entries (a power of two above 2 n) holding n
  // keys, for the caller to fill.
  void *raw_table(size_t slots, size_t n);

 private:
  struct Slot {
    uint32_t key, val;
  };
  std::vector<Slot> slots_;
  size_t n_ = 0, mask_ = 0;
  size_t slot(uint32_t key) const;
  void grow();
};

enum : uint8_t { CPF_CANCELLED = 1, CPF_STRANGULATION = 2, CPF_BCANC = 4 };

struct Rect {
  int x0, y0, x1, y1;
  bool contains(int x, int y) const { return x >= x0 && y >= y0 && x <= x1 && y <= y1; }
  bool interior(int x, int y) const { return x > x0 && y > y0 && x < x1 && y < y1; }
  bool on_boundary(int x, int y) const { return contains(x, y) && !interior(x, y); }
};

inline int cx_of(uint32_t c) { return (int16_t)(c & 0xFFFF); }
inline int cy_of(uint32_t c) { return (int16_t)(c >> 16); }
inline uint32_t cell_dim(uint32_t c) { return (c & 1) + ((c >> 16) & 1); }

struct DiscSpan {
  const uint32_t *p;
  size_t n;
};

struct Complex {
  Rect rect{}, ext{};
  std::vector<uint32_t> cell;
  std::vector<float> fn;
  std::vector<uint32_t> pair;
  std::vector<uint8_t> flags;
  std::vector<ConnSet> asc, des;
  // per-node disc lists as loaded from an archive: both empty when the
  // complex holds none per node, else one list per node
  std::vector<std::vector<uint32_t>> asc_disc, des_disc;
  // discs of surviving extrema (vertices of minima, quads of maxima) as
  // raster-ordered lists in CSR form by critical-point index
  std::vector<uint32_t> xd_off, xd_cells;
  // the other discs postMergeFillDiscs fills (the saddles' traced 1-manifolds
  // and the cells it appends to saddles of cancelled extrema) in CSR form by
  // slot 2 i + j (j = 0 asc, 1 des); empty until the manifold pass ran
  std::vector<uint32_t> sd_off, sd_cells;
  CellMap idx;
  // Arc insertions cancel_pairs may still make before failing (set by
  // simplify; see cancel_pairs).
  int64_t arc_budget = INT64_MAX;
  // Heap pops simplify may still make before failing with MSCX_ELIMIT
  // (bounded only under a work limit, see simplify).
  int64_t pop_budget = INT64_MAX;

  uint32_t size() const { return (uint32_t)cell.size(); }
  // capacity for n nodes in every per-node array
  void reserve_nodes(size_t n);
  // capacity a complex is built with: room for the few nodes merge_down
  // replicates into it, so that those do not move every array
  static size_t with_headroom(size_t n) { return n + n / 8 + 64; }
  // shallow_replicate_cp (grid_mscomplex.cpp:7-22)
  uint32_t replicate(const Complex &src, uint32_t j);

  void dump(std::vector<uint32_t> &out) const;
  // the asc (j = 0) or des (j = 1) disc of critical point i
  DiscSpan disc(uint32_t i, int j) const;
};

// GridDataset::writeout_connectivity from a worked tile.
void complex_from_tile(const mscx_tile *t, Complex &out);
// Same, from host copies of the tile's critical data (K critical cells;
// conn4 holds 4 words per cell).
void complex_build(Complex &c, const Rect &rect, const Rect &ext, uint32_t K, const uint32_t *crit,
                   const float *fn, const uint32_t *conn4, const uint32_t *pair);

void cancel_pairs(Complex &m, uint32_t cp0, uint32_t cp1, std::vector<uint32_t> *new_edges);
void uncancel_pairs(Complex &m, uint32_t cp1, uint32_t cp2);
void merge_up(const Complex &a, const Complex &b, Complex &out);
void merge_down(Complex &parent, Complex &a, Complex &b);
// Parts of the merges: the shared line of a and b (rect_t::intersection
// with the reference's checks), the line's cancellations after merge_up's
// union and its uncancellations before merge_down's copies.
void merge_ixn(const Complex &a, const Complex &b, Rect &ixn);
void merge_ixn(const Rect &ra, const Rect &rb, Rect &ixn);
void merge_rects(const Rect &r1, const Rect &e1, const Rect &r2, const Rect &e2, Rect &r, Rect &e);
void merge_up_line(Complex &m, const Rect &ixn);
void merge_down_line(Complex &P, const Rect &ixn);
// work_limit > 0 bounds the work of one call in proportion to the complex
// (no counterpart in the reference, which runs until done): simplify fails
// with MSCX_ELIMIT after min(2^27, max(2^24, work_limit x initial heap
// entries)) pops,
// and either call with MSCX_ENOMEM after max(2^24, 2 x work_limit x initial
// arc entries) inserted arc entries (at most the fixed 2^28 cap).
void simplify(Complex &m, std::vector<std::pair<uint32_t, uint32_t>> &canc, double thresh, double work_limit = 0);
void un_simplify(Complex &m, const std::vector<std::pair<uint32_t, uint32_t>> &canc, double work_limit = 0);

// GridDataset::postMergeFillDiscs from a tile's codes and owner image (host
// copies over ext, x fastest).
void fill_discs(Complex &m, const Rect &rect, const Rect &ext, const uint8_t *codes, std::vector<uint32_t> &own);

// Parts of fill_discs (the pipeline runs the label pass on the GPU).
struct CodesView {
  const uint8_t *p;
  int64_t pitch;
  int ox, oy;
  uint8_t at(int x, int y) const { return p[(int64_t)(y - oy) * pitch + (x - ox)]; }
};
struct DiscPrep {
  // (cell, owner), sorted by cell (only when built with the maps); redir
  // finds a cell's entry
  std::vector<std::pair<uint32_t, uint32_t>> redirect;
  CellMap redir;
  // per node: the node its cell's owner is redirected to (the last
  // assignment wins), UINT32_MAX when none; the redirected cells are always
  // nodes of the complex (a boundary-cancellable saddle and its pair), and
  // so are the targets (the surviving extremum each saddle is connected to)
  std::vector<uint32_t> redir_node;
  // (cell, cp index) of the surviving minima / maxima, sorted by cell (only
  // when built with the maps)
  std::vector<std::pair<uint32_t, uint32_t>> minmap, maxmap;
  std::vector<uint32_t> surv_saddles;
  // the entries postMergeFillDiscs appends to each disc before tracing, in
  // its order (CSR by slot 2 i + j), and per slot whether it is traced (the
  // discs of surviving saddles)
  std::vector<uint32_t> init_off, init_cells;
  std::vector<uint8_t> traced;
};
// One disc traced from its initial entries (the reference's loop over the
// growing list, grid_dataset.cpp:1189-1228): appends the whole list to out.
void disc_trace_one(const uint32_t *init, size_t n_init, int j, const Rect &rect, const Rect &ext,
                    const CodesView &cv, std::vector<uint32_t> &out);
// Every slot's disc (traced or copied) into m.sd_off / m.sd_cells.
void disc_trace_host(Complex &m, const DiscPrep &dp, const Rect &rect, const Rect &ext, const CodesView &cv);
// The critical-point pass.  with_maps = false (the GPU manifold pass, which
// looks surviving extrema up through the complex's own cell map) leaves the
// sorted vectors out.
void disc_prepare(Complex &m, const Rect &rect, const CodesView &cv, DiscPrep &out, bool with_maps = true);
// The surviving (not boundary-cancellable) extremum of dimension dim on
// cell c, or UINT32_MAX: what a lookup in minmap / maxmap gives.
uint32_t surviving_extremum(const Complex &m, uint32_t c, uint32_t dim);
void disc_labels_host(const DiscPrep &dp, const Rect &rect, const Rect &ext, const uint32_t *own, int64_t own_pitch,
                      std::vector<uint32_t> &lv, std::vector<uint32_t> &lq);
void disc_group(Complex &m, const Rect &rect, const uint32_t *lv, const uint32_t *lq);


// The gradient codes of n worked tiles copied to pinned host memory ahead
// of their manifold pass (discs.cpp); the copies run on the auxiliary stream.
struct CodesStage;
std::shared_ptr<CodesStage> stage_codes(mscx_tile *const *tiles, int n);
// postMergeFillDiscs for n worked tiles and their complexes, label pass on
// the GPU (discs.cpp); `pre` (may be null) holds the same tiles' codes
// staged by stage_codes.
void fill_discs_device(mscx_tile *const *tiles, Complex *const *cs, int n, CodesStage *pre = nullptr);

void write_discs(const Complex &m, const std::string &prefix);
// Host archive of a complex (dump words + checksum); load throws MSCX_EIO
// on a missing, truncated or corrupt file.
void save_complex(const Complex &m, const std::string &path);
void load_complex(const std::string &path, Complex &m);
// A complex from its dump words (the body of an archive); MSCX_EIO when they
// are malformed (`what` names their source in the message).
void complex_from_words(const uint32_t *w, size_t n, Complex &m, const std::string &what);

// A complex set aside by the out-of-core mode: its nodes as records in the
// dump's per-node format (the discs there are only the per-node vectors) in
// chunks that are written and read on several threads, its cell map and
// its disc CSRs moved as they are.  spill_complex leaves m empty;
// restore_complex rebuilds it exactly and leaves `in` empty.
struct ComplexSpill {
  Rect rect{}, ext{};
  uint32_t n = 0;
  bool node_discs = false;  // the complex held per-node disc lists
  std::vector<uint32_t> begin;  // first node of each chunk, then n
  std::vector<std::vector<uint32_t>> words;
  CellMap idx;
  std::vector<uint32_t> xd_off, xd_cells, sd_off, sd_cells;
};
void spill_complex(Complex &m, ComplexSpill &out);
void restore_complex(ComplexSpill &in, Complex &m);

}  // namespace mscx

// Opaque C handle of a complex.
struct mscx_complex {
  mscx::Complex c;
};
