// CDNA4 (gfx950) kernels for the per-tile discrete Morse-Smale stages:
// gradient assignment, critical-cell collation (segment counts, scan,
// compaction), manifold tracing (pointer jumping), owner images and saddle
// incidence.  All kernels take an array of tiles so one launch covers every
// tile of a tiling or of a batch of grids.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mscx {

__device__ __forceinline__ int find_tile(const int *__restrict__ prefix, int ntiles, int b) {
  int lo = 0, hi = ntiles - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ uint32_t pack_cell(int x, int y) { return (uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16); }

// ------------------------------------------------------------------ gradient
// One block = GRAD_TX x GRAD_TY vertices (4 waves, one vertex row each, two
// vertices per lane).  Vertex keys of the tile plus a 1/2 vertex halo are
// staged in LDS; each lane emits the codes of its 2x2 cells per vertex, so a
// wave writes one 256-cell collation segment of two cell rows and the
// critical counts of both.  The successor of every vertex and quad (the
// trace's starting pointers) is written here too.
__global__ __launch_bounds__(256) void k_gradient(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                  int ntiles) {
  constexpr int LW = GRAD_TX + 3, LH = GRAD_TY + 3;
  __shared__ uint32_t sk[LW * LH];
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc &T = tiles[t];
  const TileGeom g = T.g;
  const int lb = b - prefix[t];
  const int nbx = (g.vx + GRAD_TX - 1) / GRAD_TX;
  const int bx = lb % nbx, by = lb / nbx;
  const int i0 = bx * GRAD_TX, j0 = by * GRAD_TY;

  for (int idx = threadIdx.x; idx < LW * LH; idx += 256) {
    int li = idx % LW, lj = idx / LW;
    int gi = i0 - 1 + li, gj = j0 - 1 + lj;
    uint32_t k = 0;
    if (gi >= 0 && gj >= 0 && gi < g.vx && gj < g.vy) k = fkey(T.verts[(int64_t)gj * T.vpitch + gi]);
    sk[idx] = k;
  }
  __syncthreads();

  auto K = [&](int i, int j) -> uint64_t { return vkey(sk[(j - j0 + 1) * LW + (i - i0 + 1)], i, j); };

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = j0 + w;
  if (j >= g.vy) return;
  const int ia = i0 + 2 * lane;
  uint32_t c0 = 0, c1 = 0, n0 = 0, n1 = 0;
  const bool row1 = (2 * j + 1) < g.ch;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int cx = 2 * ia + k;
    if (cx < g.cw) {
      uint8_t c = cell_code(K, g, cx, 2 * j);
      c0 |= (uint32_t)c << (8 * k);
      n0 += (c >> 1) & 1;
      if (row1) {
        c = cell_code(K, g, cx, 2 * j + 1);
        c1 |= (uint32_t)c << (8 * k);
        n1 += (c >> 1) & 1;
      }
    }
  }
  uint8_t *row = T.codes + (int64_t)(2 * j) * T.pitch + 2 * i0 + 4 * lane;
  *(uint32_t *)row = c0;
  if (row1) *(uint32_t *)(row + T.pitch) = c1;

  // successors for the trace
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    int i = ia + k;
    if (i < g.vx) {
      uint8_t cv = (uint8_t)(c0 >> (16 * k));
      int32_t p = -1;
      if ((cv & 3) == CF_PAIRED) {
        int d = code_dir(cv);
        p = (j + dir_dy(d)) * g.vx + (i + dir_dx(d));
      }
      T.vptr[(int64_t)j * g.vx + i] = p;
      if (row1 && i + 1 < g.vx) {
        uint8_t cq = (uint8_t)(c1 >> (16 * k + 8));
        int32_t q = -1;
        if ((cq & 3) == CF_PAIRED) {
          int d = code_dir(cq);
          q = (j + dir_dy(d)) * (g.vx - 1) + (i + dir_dx(d));
        }
        T.qptr[(int64_t)j * (g.vx - 1) + i] = q;
      }
    }
  }

  uint32_t s0 = wave_sum(n0), s1 = wave_sum(n1);
  if (lane == 0) {
    T.seg_counts[(int64_t)(2 * j) * T.nseg + bx] = s0;
    if (row1) T.seg_counts[(int64_t)(2 * j + 1) * T.nseg + bx] = s1;
  }
}

// ---------------------------------------------------------------------- scan
// Exclusive scan of a tile's (row, segment) counts in raster order, in three
// launches: chunk sums, per-tile scan of the chunk sums (writes K), and the
// chunk-local scans.

__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t *s_wave, uint32_t &total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
  for (int q = 0; q < 4; ++q) {
    if (q < w) wbase += s_wave[q];
    tot += s_wave[q];
  }
  __syncthreads();
  total = tot;
  return wbase + x - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                     int ntiles) {
  __shared__ uint32_t s_wave[4];
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc &T = tiles[t];
  const int lb = b - prefix[t];
  const int64_t n = (int64_t)T.g.ch * T.nseg;
  const int64_t base = (int64_t)lb * SCAN_CHUNK;
  uint32_t s = 0;
  for (int k = 0; k < SCAN_CHUNK / 256; ++k) {
    int64_t i = base + k * 256 + threadIdx.x;
    if (i < n) s += T.seg_counts[i];
  }
  uint32_t tot;
  block_exclusive_scan_256(s, s_wave, tot);
  if (threadIdx.x == 0) T.scan_part[lb] = tot;
}

__global__ __launch_bounds__(256) void k_scan_parts(const TileDesc *__restrict__ tiles) {
  __shared__ uint32_t s_wave[4];
  const TileDesc &T = tiles[blockIdx.x];
  const int64_t n = (int64_t)T.g.ch * T.nseg;
  const int nparts = (int)((n + SCAN_CHUNK - 1) / SCAN_CHUNK);
  uint32_t carry = 0;
  for (int base = 0; base < nparts; base += 256) {
    int i = base + threadIdx.x;
    uint32_t v = (i < nparts) ? T.scan_part[i] : 0;
    uint32_t tot;
    uint32_t ex = block_exclusive_scan_256(v, s_wave, tot);
    if (i < nparts) T.scan_part[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) T.nkrit[0] = carry;
}

__global__ __launch_bounds__(256) void k_scan_down(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                   int ntiles) {
  __shared__ uint32_t s_wave[4];
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc &T = tiles[t];
  const int lb = b - prefix[t];
  const int64_t n = (int64_t)T.g.ch * T.nseg;
  const int64_t base = (int64_t)lb * SCAN_CHUNK + (int64_t)threadIdx.x * (SCAN_CHUNK / 256);
  uint32_t v[SCAN_CHUNK / 256];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_CHUNK / 256; ++k) {
    int64_t i = base + k;
    v[k] = (i < n) ? T.seg_counts[i] : 0;
    s += v[k];
  }
  uint32_t tot;
  uint32_t ex = block_exclusive_scan_256(s, s_wave, tot) + T.scan_part[lb];
#pragma unroll
  for (int k = 0; k < SCAN_CHUNK / 256; ++k) {
    int64_t i = base + k;
    if (i < n) T.seg_offs[i] = ex;
    ex += v[k];
  }
}

// ---------------------------------------------------------------- compaction
// One wave per (cell row, segment): critical cells are written in raster
// order at their scanned offsets with their cell function value
// (grid_dataset.cpp:1538-1552: float sum in getCellPoints order, then the
// division), and critical vertices/quads become resolved trace roots.

__device__ __forceinline__ float vval(const TileDesc &T, int i, int j) { return T.verts[(int64_t)j * T.vpitch + i]; }

__device__ __forceinline__ float cell_fn(const TileDesc &T, int cx, int cy) {
  int i = cx >> 1, j = cy >> 1;
  int dx = cx & 1, dy = cy & 1;
  float fn = 0.0f;
  if (!dx && !dy) {
    fn += vval(T, i, j);
    return fn / 1.0f;
  }
  if (dx && !dy) {  // points (x+1,y), (x-1,y)
    fn += vval(T, i + 1, j);
    fn += vval(T, i, j);
    return fn / 2.0f;
  }
  if (!dx && dy) {  // points (x,y+1), (x,y-1)
    fn += vval(T, i, j + 1);
    fn += vval(T, i, j);
    return fn / 2.0f;
  }
  // points (x+1,y+1), (x+1,y-1), (x-1,y-1), (x-1,y+1)
  fn += vval(T, i + 1, j + 1);
  fn += vval(T, i + 1, j);
  fn += vval(T, i, j);
  fn += vval(T, i, j + 1);
  return fn / 4.0f;
}

__global__ __launch_bounds__(256) void k_compact(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                 int ntiles) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc &T = tiles[t];
  const TileGeom g = T.g;
  const int lb = b - prefix[t];
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)lb * 4 + (threadIdx.x >> 6);
  if (wid >= (int64_t)g.ch * T.nseg) return;
  const int row = (int)(wid / T.nseg), seg = (int)(wid % T.nseg);
  const uint32_t word = *(const uint32_t *)(T.codes + (int64_t)row * T.pitch + seg * SEG + 4 * lane);
  uint32_t base = T.seg_offs[wid];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t before = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint64_t m = __ballot((word >> (8 * k + 1)) & 1);
    before += __popcll(m & lt);
  }
  uint32_t idx = base + before;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint8_t c = (uint8_t)(word >> (8 * k));
    if (c & CF_CRITICAL) {
      int cx = seg * SEG + 4 * lane + k, cy = row;
      if (idx < T.kcap) {
        T.crit[idx] = pack_cell(T.ox + cx, T.oy + cy);
        T.crit_fn[idx] = cell_fn(T, cx, cy);
        if (!(cx & 1) && !(cy & 1)) T.vptr[(int64_t)(cy >> 1) * g.vx + (cx >> 1)] = -(int32_t)idx - 2;
        if ((cx & 1) && (cy & 1)) T.qptr[(int64_t)(cy >> 1) * (g.vx - 1) + (cx >> 1)] = -(int32_t)idx - 2;
      }
      ++idx;
    }
  }
}

// --------------------------------------------------------------------- trace
// One pointer-jumping round over every vertex and quad pointer of every
// tile: p <- ptr[p] while p names another vertex/quad.  Rounds repeat until
// no pointer is unresolved (the host checks *changed every few rounds).

__global__ __launch_bounds__(256) void k_jump(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                              int ntiles, uint32_t *changed) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc &T = tiles[t];
  const int lb = b - prefix[t];
  const int64_t nv = (int64_t)T.g.vx * T.g.vy, nq = (int64_t)(T.g.vx - 1) * (T.g.vy - 1);
  bool any = false;
  for (int k = 0; k < 4; ++k) {
    int64_t e = ((int64_t)lb * 4 + k) * 256 + threadIdx.x;
    int32_t *arr;
    int64_t i;
    if (e < nv) {
      arr = T.vptr;
      i = e;
    } else if (e < nv + nq) {
      arr = T.qptr;
      i = e - nv;
    } else {
      break;
    }
    int32_t p = arr[i];
    if (p >= 0) {
      int32_t q = arr[p];
      arr[i] = q;
      any |= (q >= 0);
    }
  }
  if (__any(any) && (threadIdx.x & 63) == 0) atomicOr(changed, 1u);
}

// ---------------------------------------------------------------- owners
// Owner image (the reference's m_cell_own): critical cells own themselves;
// vertices and quads take their resolved root; an edge paired with a vertex
// (quad) takes that vertex's (quad's) owner; cells never reached stay ~0.

__device__ __forceinline__ uint32_t own_of_ptr(const TileDesc &T, int32_t p) {
  return (p <= -2) ? T.crit[-p - 2] : 0xFFFFFFFFu;
}

__global__ __launch_bounds__(256) void k_owner(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                               int ntiles) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc &T = tiles[t];
  const TileGeom g = T.g;
  const int lb = b - prefix[t];
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)lb * 4 + (threadIdx.x >> 6);
  if (wid >= (int64_t)g.ch * T.nseg) return;
  const int row = (int)(wid / T.nseg), seg = (int)(wid % T.nseg);
  const uint32_t word = *(const uint32_t *)(T.codes + (int64_t)row * T.pitch + seg * SEG + 4 * lane);
  uint32_t o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint8_t c = (uint8_t)(word >> (8 * k));
    int cx = seg * SEG + 4 * lane + k, cy = row;
    uint32_t v = 0xFFFFFFFFu;
    if (cx < g.cw) {
      int dx = cx & 1, dy = cy & 1;
      if (dx == dy) {  // vertex or quad
        v = dx ? own_of_ptr(T, T.qptr[(int64_t)(cy >> 1) * (g.vx - 1) + (cx >> 1)])
               : own_of_ptr(T, T.vptr[(int64_t)(cy >> 1) * g.vx + (cx >> 1)]);
      } else if (c & CF_CRITICAL) {
        v = pack_cell(T.ox + cx, T.oy + cy);
      } else if (c & CF_PAIRED) {
        int d = code_dir(c);
        int px = cx + dir_dx(d), py = cy + dir_dy(d);
        v = (px & 1) ? own_of_ptr(T, T.qptr[(int64_t)(py >> 1) * (g.vx - 1) + (px >> 1)])
                     : own_of_ptr(T, T.vptr[(int64_t)(py >> 1) * g.vx + (px >> 1)]);
      }
    }
    o[k] = v;
  }
  uint4 *dst = (uint4 *)(T.own + (int64_t)row * T.own_pitch + seg * SEG + 4 * lane);
  *dst = make_uint4(o[0], o[1], o[2], o[3]);
}

// --------------------------------------------------------------- incidence
// Per critical cell: for saddles, the critical index owning each facet and
// each in-ext cofacet (writeout_connectivity, grid_dataset.cpp:1876-1898);
// for boundary-cancellable cells, the critical index of the pair.

__device__ __forceinline__ uint32_t idx_of_ptr(int32_t p) { return (p <= -2) ? (uint32_t)(-p - 2) : 0xFFFFFFFFu; }

__device__ uint32_t find_crit(const TileDesc &T, uint32_t K, int gx, int gy) {
  // packed cellids (x | y << 16) sort numerically in raster (y, x) order
  const uint32_t key = pack_cell(gx, gy);
  uint32_t lo = 0, hi = K;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (T.crit[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < K && T.crit[lo] == key) ? lo : 0xFFFFFFFFu;
}

__global__ __launch_bounds__(256) void k_incidence(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                   int ntiles) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc &T = tiles[t];
  const TileGeom g = T.g;
  const int lb = b - prefix[t];
  const uint32_t K = T.nkrit[0];
  const uint32_t k = (uint32_t)lb * 256 + threadIdx.x;
  if (k >= K || k >= T.kcap) return;
  const uint32_t c = T.crit[k];
  const int cx = (int)(c & 0xFFFF) - T.ox, cy = (int)(c >> 16) - T.oy;
  const uint8_t code = T.codes[(int64_t)cy * T.pitch + cx];
  const int dx = cx & 1, dy = cy & 1;
  uint4 conn = make_uint4(~0u, ~0u, ~0u, ~0u);
  if (dx != dy) {
    // facets: the two vertices; cofacets: the two quads across the edge
    int fx0 = cx + dx, fy0 = cy + dy, fx1 = cx - dx, fy1 = cy - dy;
    conn.x = idx_of_ptr(T.vptr[(int64_t)(fy0 >> 1) * g.vx + (fx0 >> 1)]);
    conn.y = idx_of_ptr(T.vptr[(int64_t)(fy1 >> 1) * g.vx + (fx1 >> 1)]);
    int qx0 = cx + dy, qy0 = cy + dx, qx1 = cx - dy, qy1 = cy - dx;
    if (qx0 >= 0 && qy0 >= 0 && qx0 < g.cw && qy0 < g.ch)
      conn.z = idx_of_ptr(T.qptr[(int64_t)(qy0 >> 1) * (g.vx - 1) + (qx0 >> 1)]);
    if (qx1 >= 0 && qy1 >= 0 && qx1 < g.cw && qy1 < g.ch)
      conn.w = idx_of_ptr(T.qptr[(int64_t)(qy1 >> 1) * (g.vx - 1) + (qx1 >> 1)]);
  }
  ((uint4 *)T.crit_conn)[k] = conn;
  uint32_t pidx = ~0u;
  if (code & CF_PAIRED) {
    int d = code_dir(code);
    int px = cx + dir_dx(d), py = cy + dir_dy(d);
    int pdx = px & 1, pdy = py & 1;
    if (pdx == pdy)
      pidx = pdx ? idx_of_ptr(T.qptr[(int64_t)(py >> 1) * (g.vx - 1) + (px >> 1)])
                 : idx_of_ptr(T.vptr[(int64_t)(py >> 1) * g.vx + (px >> 1)]);
    else
      pidx = find_crit(T, K, T.ox + px, T.oy + py);
  }
  T.crit_pair[k] = pidx;
}

// ------------------------------------------------------------ disc labels
// For every interior vertex (quad) of a leaf: o = owner after redirects; if
// o lies in the rect, one more redirect step o = owner(o); the label is the
// critical-point index of surviving extremum o.

__device__ __forceinline__ uint32_t table_find(const uint32_t *__restrict__ t, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (t[2 * mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < n && t[2 * lo] == key) ? t[2 * lo + 1] : 0xFFFFFFFFu;
}

__device__ __forceinline__ uint32_t owner_after(const DiscDesc &D, int x, int y) {
  uint32_t r = D.nredir ? table_find(D.redir, D.nredir, pack_cell(x, y)) : 0xFFFFFFFFu;
  if (r != 0xFFFFFFFFu) return r;
  return D.own[(int64_t)(y - D.oy) * D.own_pitch + (x - D.ox)];
}

__global__ __launch_bounds__(256) void k_disc_labels(const DiscDesc *__restrict__ descs, const int *__restrict__ prefix,
                                                     int n) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, n, b);
  const DiscDesc &D = descs[t];
  const int nvx = (D.rx1 - D.rx0) / 2 + 1, nvy = (D.ry1 - D.ry0) / 2 + 1;
  const int64_t nv = (int64_t)nvx * nvy, nq = (int64_t)(nvx - 1) * (nvy - 1);
  const int64_t item = (int64_t)(b - prefix[t]) * 256 + threadIdx.x;
  if (item >= nv + nq) return;
  const bool vert = item < nv;
  const int64_t r = vert ? item : item - nv;
  const int w = vert ? nvx : nvx - 1;
  const int x = D.rx0 + (vert ? 0 : 1) + 2 * (int)(r % w), y = D.ry0 + (vert ? 0 : 1) + 2 * (int)(r / w);
  uint32_t o = owner_after(D, x, y);
  if (o != 0xFFFFFFFFu) {
    int ox = (int)(int16_t)(o & 0xFFFF), oy = (int)(int16_t)(o >> 16);
    if (ox >= D.rx0 && ox <= D.rx1 && oy >= D.ry0 && oy <= D.ry1) o = owner_after(D, ox, oy);
  }
  uint32_t lab = 0xFFFFFFFFu;
  if ((o & 0xFFFF) != 0xFFFF && (o >> 16) != 0xFFFF) {
    lab = vert ? table_find(D.minmap, D.nmin, o) : table_find(D.maxmap, D.nmax, o);
    if (lab == 0xFFFFFFFFu) atomicOr(D.err, 1u);
  }
  (vert ? D.lab_v : D.lab_q)[r] = lab;
}

// -------------------------------------------------------------- launchers

static int blocks_of(int kind, const TileDesc &T) {
  const TileGeom &g = T.g;
  switch (kind) {
    case KB_GRADIENT:
      return ((g.vx + GRAD_TX - 1) / GRAD_TX) * ((g.vy + GRAD_TY - 1) / GRAD_TY);
    case KB_SCAN: {
      int64_t n = (int64_t)g.ch * T.nseg;
      return (int)((n + SCAN_CHUNK - 1) / SCAN_CHUNK);
    }
    case KB_ROWSEG: {
      int64_t n = (int64_t)g.ch * T.nseg;
      return (int)((n + 3) / 4);
    }
    case KB_JUMP: {
      int64_t n = (int64_t)g.vx * g.vy + (int64_t)(g.vx - 1) * (g.vy - 1);
      return (int)((n + 1023) / 1024);
    }
    case KB_CRIT:
      return (int)((T.kcap + 255) / 256);
  }
  return 0;
}

int plan_blocks(int kind, const TileDesc *tiles, int ntiles, int *prefix) {
  int acc = 0;
  for (int t = 0; t < ntiles; ++t) {
    prefix[t] = acc;
    acc += blocks_of(kind, tiles[t]);
  }
  prefix[ntiles] = acc;
  return acc;
}

int plan_disc_blocks(const DiscDesc *d, int n, int *prefix) {
  int acc = 0;
  for (int t = 0; t < n; ++t) {
    prefix[t] = acc;
    int64_t nvx = (d[t].rx1 - d[t].rx0) / 2 + 1, nvy = (d[t].ry1 - d[t].ry0) / 2 + 1;
    int64_t items = nvx * nvy + (nvx - 1) * (nvy - 1);
    acc += (int)((items + 255) / 256);
  }
  prefix[n] = acc;
  return acc;
}

void launch_disc_labels(const DiscDesc *d_desc, const int *d_prefix, int n, int nblocks, hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_disc_labels, dim3(nblocks), dim3(256), 0, s, d_desc, d_prefix, n);
}

void launch_gradient(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_gradient, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles);
}
void launch_scan(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s) {
  if (!nblocks) return;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles);
  hipLaunchKernelGGL(k_scan_parts, dim3(ntiles), dim3(256), 0, s, d_tiles);
  hipLaunchKernelGGL(k_scan_down, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles);
}
void launch_compact(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_compact, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles);
}
void launch_jump(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, uint32_t *d_changed,
                 hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_jump, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles, d_changed);
}
void launch_owner(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_owner, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles);
}
void launch_incidence(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_incidence, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles);
}

}  // namespace mscx
