// mscx_dmsc: the reference's dmsc command line over the C ABI
// (main.cpp:106-205 parse_grid, the GridDataManager run of
// grid_datamanager.cpp:596-662).
//
//   mscx_dmsc -gf <filename> -d <x_dim> <y_dim> [-l <levels>] [-st] [-ocl]
//             [--simp-tresh 0.<tresh>] [--out-of-core] [--num-parallel <num>]
//
// Flags keep the reference's meaning and defaults: levels 1 (2 slabs along
// y), num_parallel 1 (which makes the run single-threaded, as
// GridDataManager's constructor does), threshold 0 (no simplification);
// -st forces one host thread.  Every piece's discs are appended to
// dp__disc_asc_(x,y) / dp__disc_des_(x,y) in the working directory, as
// collectManifold writes them (grid_datamanager.cpp:397-420).  The GPU path
// always runs, so -ocl is accepted and changes nothing; --out-of-core keeps
// the pieces of the merge tree that no phase is using serialised
// (MSCX_RUN_OUT_OF_CORE; archive files under $MSCX_SPILL_DIR when it is
// set), as the reference spills them to archive files.  The reference matches its flags with regular
// expressions over the joined command line and ignores what does not
// match; here each flag takes its value from the next argument(s), a file
// name is taken whole, and an unknown argument or a malformed value is a
// usage error.  Additions: --lx/--ly select a 2-D tiling, --no-write skips
// the disc files, --print-args prints the parsed settings and exits without
// touching the GPU.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/mscx.h"

namespace {

struct Args {
  std::string file;
  long sx = 0, sy = 0;
  long levels = 1, lx = 0, ly = 0;
  bool single_thread = false, ocl = false, out_of_core = false, write = true, print = false;
  double thresh = 0.0;
  long num_parallel = 1;
};

void usage(const char *prog) {
  fprintf(stderr,
          "Usage: %s -gf <filename> -d <x_dim> <y_dim> [-l <levels>] [-st] [-ocl] [--simp-tresh 0.<tresh>]"
          " [--out-of-core] [--num-parallel <num>] [--lx <n> --ly <n>] [--no-write] [--print-args]\n",
          prog);
}

bool to_long(const char *s, long lo, long hi, long &out) {
  if (!s || !*s) return false;
  char *end = nullptr;
  errno = 0;
  const long v = strtol(s, &end, 10);
  if (errno || *end || v < lo || v > hi) return false;
  out = v;
  return true;
}

bool to_double(const char *s, double &out) {
  if (!s || !*s) return false;
  char *end = nullptr;
  errno = 0;
  const double v = strtod(s, &end);
  if (errno || *end || !(v >= 0.0 && v < 1.0)) return false;  // the reference accepts 0.<digits> only
  out = v;
  return true;
}

bool parse(int argc, char **argv, Args &a) {
  bool have_file = false, have_dim = false;
  for (int i = 1; i < argc; ++i) {
    const std::string f = argv[i];
    auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
    if (f == "-gf") {
      const char *v = next();
      if (!v) return false;
      a.file = v;
      have_file = true;
    } else if (f == "-d") {
      if (!to_long(next(), 2, 16384, a.sx) || !to_long(next(), 2, 16384, a.sy)) return false;
      have_dim = true;
    } else if (f == "-l") {
      if (!to_long(next(), 0, 12, a.levels)) return false;
    } else if (f == "--lx") {
      if (!to_long(next(), 0, 12, a.lx)) return false;
    } else if (f == "--ly") {
      if (!to_long(next(), 0, 12, a.ly)) return false;
    } else if (f == "--num-parallel") {
      if (!to_long(next(), 1, 1024, a.num_parallel)) return false;
    } else if (f == "--simp-tresh") {
      if (!to_double(next(), a.thresh)) return false;
    } else if (f == "-st") {
      a.single_thread = true;
    } else if (f == "-ocl") {
      a.ocl = true;
    } else if (f == "--out-of-core") {
      a.out_of_core = true;
    } else if (f == "--no-write") {
      a.write = false;
    } else if (f == "--print-args") {
      a.print = true;
    } else {
      return false;
    }
  }
  if (a.single_thread) a.num_parallel = 1;  // GridDataManager: -st means one thread
  return have_file && have_dim;
}

}  // namespace

int main(int argc, char **argv) {
  Args a;
  if (!parse(argc, argv, a)) {
    usage(argv[0]);
    return 2;
  }
  if (a.print) {
    printf("{\"file\": \"%s\", \"size\": [%ld, %ld], \"levels\": %ld, \"lx\": %ld, \"ly\": %ld, "
           "\"simp_tresh\": %.17g, \"num_parallel\": %ld, \"single_threaded\": %s, \"write_discs\": %s, "
           "\"out_of_core\": %s}\n",
           a.file.c_str(), a.sx, a.sy, a.levels, a.lx, a.ly, a.thresh, a.num_parallel,
           a.num_parallel == 1 ? "true" : "false", a.write ? "true" : "false", a.out_of_core ? "true" : "false");
    return 0;
  }
  mscx_grid *g = nullptr;
  int rc = mscx_grid_run_file(a.file.c_str(), (int32_t)a.sx, (int32_t)a.sy, (int32_t)a.levels, (int32_t)a.lx,
                              (int32_t)a.ly, a.thresh, (int32_t)a.num_parallel, a.out_of_core ? MSCX_RUN_OUT_OF_CORE : 0u,
                              a.write ? "dp__disc_" : nullptr, &g);
  if (rc) {
    fprintf(stderr, "mscx_dmsc: error %d: %s\n", rc, mscx_last_error());
    return 1;
  }
  int32_t n = 0;
  double t[5] = {0, 0, 0, 0, 0};
  const mscx_complex *root = nullptr;
  uint32_t ncp = 0;
  rc = mscx_grid_num_pieces(g, &n);
  if (!rc) rc = mscx_grid_times(g, t);
  if (!rc) rc = mscx_grid_piece(g, n - 1, &root);
  if (!rc) rc = mscx_complex_num_cps(root, &ncp);
  if (rc) {
    fprintf(stderr, "mscx_dmsc: error %d: %s\n", rc, mscx_last_error());
    mscx_grid_destroy(g);
    return 1;
  }
  const double total = t[0] + t[1] + t[2] + t[3] + t[4];
  printf("pieces %d, root critical points %u\n", n, ncp);
  printf("read+compute %.3f s, merge up %.3f s, simplify %.3f s, merge down %.3f s, discs %.3f s\n", t[0], t[1], t[2],
         t[3], t[4]);
  printf("total %.3f s, %.2f Mcells/s\n", total, (double)a.sx * (double)a.sy / 1e6 / (total > 0 ? total : 1e-9));
  mscx_grid_destroy(g);
  mscx_shutdown();
  return 0;
}
