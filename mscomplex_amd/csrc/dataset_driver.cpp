// mscx_dataset_driver: the reference's GridDataManager compute order
// (grid_datamanager.cpp:40-89, 121-210, 397-451, 596-662) written against
// the header-only drop-in include/mscx_dataset.hpp, as a dmsc maintainer
// would call it: slab pieces by createPieces_quadtree's split rule, work()
// + writeout_connectivity() per piece, merge_up by pairs, the root's
// simplify_un_simplify, merge_down parents first, then postMergeFillDiscs per
// leaf.
//
//   mscx_dataset_driver <grid.raw> <size_x> <size_y> <levels> <threshold> <out.bin> [<arrays.bin>]
//
// out.bin: the number of pieces, then per piece (leaves first, root last)
// its dump length and dump words.  arrays.bin (optional): leaf 0's
// m_cell_flags, m_cell_pairs and m_cell_own in the reference's [x][y]
// layout, then m_critical_cells.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <vector>

#include "../../include/mscx_dataset.hpp"

using namespace mscx_dmsc;

namespace {

struct Piece {
  rect_t r, e;
  std::unique_ptr<GridDataset> ds;
  std::unique_ptr<GridMSComplex> msc;
};

// createPieces_quadtree (grid_datamanager.cpp:53-89): split along y only;
// the children share the split row and grow their ext by one vertex past it
void create_pieces(const rect_t &r, const rect_t &e, int level, std::vector<Piece> &out) {
  if (level == 0) {
    out.push_back(Piece{r, e, nullptr, nullptr});
    return;
  }
  const int sy = r.top() - r.bottom();
  const int split = r.top() - 2 * (sy / 4);
  create_pieces(rect_t(r.left(), r.bottom(), r.right(), split), rect_t(e.left(), e.bottom(), e.right(), split + 2),
                level - 1, out);
  create_pieces(rect_t(r.left(), split, r.right(), r.top()), rect_t(e.left(), split - 2, e.right(), e.top()),
                level - 1, out);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 7) {
    fprintf(stderr, "usage: %s <grid.raw> <size_x> <size_y> <levels> <threshold> <out.bin> [<arrays.bin>]\n", argv[0]);
    return 2;
  }
  try {
    const int sx = atoi(argv[2]), sy = atoi(argv[3]), levels = atoi(argv[4]);
    const double thresh = atof(argv[5]);
    std::vector<float> grid((size_t)sx * sy);
    {
      std::ifstream is(argv[1], std::ios::binary);
      is.read((char *)grid.data(), (std::streamsize)(grid.size() * 4));
      if (!is) throw std::runtime_error("unable to read stream");
    }
    GridDataset::init_opencl();
    const rect_t whole(0, 0, 2 * (sx - 1), 2 * (sy - 1));
    std::vector<Piece> pieces;
    create_pieces(whole, whole, levels, pieces);
    const size_t nleaf = pieces.size();
    // computeMsGraph (:121-140): init with the ext window, work, writeout
    for (Piece &p : pieces) {
      p.ds.reset(new GridDataset(p.r, p.e));
      p.msc.reset(new GridMSComplex(p.r, p.e));
      const int x0 = p.e.left() / 2, y0 = p.e.bottom() / 2;
      const int w = p.ds->num_vert(0), h = p.ds->num_vert(1);
      std::vector<float> win((size_t)w * h);
      for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) win[(size_t)y * w + x] = grid[(size_t)(y0 + y) * sx + (x0 + x)];
      p.ds->init(win.data());
      p.ds->work();
      p.ds->writeout_connectivity(p.msc.get());
    }
    if (argc > 7) {
      const GridDataset &d = *pieces[0].ds;
      std::ofstream os(argv[7], std::ios::binary);
      os.write((const char *)d.m_cell_flags.data(), (std::streamsize)d.m_cell_flags.size());
      os.write((const char *)d.m_cell_pairs.data(), (std::streamsize)(d.m_cell_pairs.size() * 4));
      os.write((const char *)d.m_cell_own.data(), (std::streamsize)(d.m_cell_own.size() * 4));
      os.write((const char *)d.m_critical_cells.data(), (std::streamsize)(d.m_critical_cells.size() * 4));
    }
    // mergePiecesUp (:258-306): pairs (j, j+1) of a level, parent appended
    for (size_t lo = 0, width = nleaf; width > 1; lo += width, width /= 2)
      for (size_t k = 0; k < width / 2; ++k) {
        Piece &a = pieces[lo + 2 * k], &b = pieces[lo + 2 * k + 1];
        Piece p;
        p.msc.reset(GridMSComplex::merge_up(*a.msc, *b.msc));
        p.r = p.msc->m_rect;
        p.e = p.msc->m_ext_rect;
        pieces.push_back(std::move(p));
      }
    // the root's simplification (:651-652)
    if (thresh > 0) pieces.back().msc->simplify_un_simplify(thresh);
    // mergePiecesDown (:308-357) and finalMergeDownPiecesInRange (:359-395):
    // parent j's children are pieces[2 (j - nleaf)] and the next one
    for (size_t j = pieces.size(); j-- > nleaf;) {
      const size_t c = 2 * (j - nleaf);
      pieces[j].msc->merge_down(*pieces[c].msc, *pieces[c + 1].msc);
    }
    // collectManifold (:397-422)
    for (size_t i = 0; i < nleaf; ++i) pieces[i].ds->postMergeFillDiscs(pieces[i].msc.get());
    std::ofstream os(argv[6], std::ios::binary);
    const uint32_t n = (uint32_t)pieces.size();
    os.write((const char *)&n, 4);
    for (const Piece &p : pieces) {
      const std::vector<uint32_t> w = p.msc->dump();
      const uint64_t len = w.size();
      os.write((const char *)&len, 8);
      os.write((const char *)w.data(), (std::streamsize)(w.size() * 4));
    }
    GridDataset::stop_opencl();
    return 0;
  } catch (const std::exception &e) {
    fprintf(stderr, "mscx_dataset_driver: %s\n", e.what());
    return 1;
  }
}
