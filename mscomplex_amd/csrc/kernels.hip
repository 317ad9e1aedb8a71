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

// ------------------------------------------------------- gradient + trace
// One block = a GRAD_TX x GRAD_TY vertex tile (cells 2*GRAD_TX x 2*GRAD_TY).
//  1. vertex keys of the tile and a 1/2 vertex halo are staged in LDS
//  2. the candidate pair of every vertex, horizontal and vertical edge the
//     tile's cells refer to is computed once (vertex_cand / hedge_cand /
//     vedge_cand) into LDS
//  3. every cell's gradient code (code_from_cands) goes to LDS
//  4. the gradient forests are traced inside the tile: each vertex points
//     across its paired edge to the next vertex and each quad across its
//     paired edge to the next quad (grid_dataset.cpp:1376-1427); pointer
//     jumping in LDS resolves every path to its critical root, to "none", or
//     to the first cell outside the tile (a marker)
//  5. codes, the owner image (edges take the owner of their paired vertex or
//     quad), per-segment critical counts and the list of marker entries are
//     written with coalesced row stores.
// A lane owns one vertex column (two cells per cell row).

__device__ __forceinline__ bool lds_is_link(uint32_t w) { return (w & 0xFFFF8000u) == 0x00008000u; }

__global__ __launch_bounds__(256) void k_grad_trace(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                    int ntiles) {
  constexpr int TX = GRAD_TX, TY = GRAD_TY;
  constexpr int KW = TX + 3, KH = TY + 3;  // keys of vertices [i0-1, i0+TX+1] x [j0-1, j0+TY+1]
  constexpr int VW = TX + 1, VH = TY + 1;  // vertex candidates [i0, i0+TX] x [j0, j0+TY]
  constexpr int HW = TX, HH = TY + 1;      // horizontal-edge candidates [i0, i0+TX) x [j0, j0+TY]
  constexpr int EW = TX + 1, EH = TY;      // vertical-edge candidates [i0, i0+TX] x [j0, j0+TY)
  constexpr int CW = 2 * TX, CH = 2 * TY;  // cells of the tile
  constexpr int NV = TX * TY;              // trace entries: vertices, then quads
  __shared__ uint32_t sk[KW * KH];
  __shared__ uint8_t svc[VW * VH], shc[HW * HH], sec[EW * EH];
  __shared__ uint8_t scode[CW * CH];
  __shared__ uint32_t sl[2 * NV];
  __shared__ uint32_t s_nexit;

  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc &T = tiles[t];
  const TileGeom g = T.g;
  const int lb = b - prefix[t];
  const int nbx = (g.vx + TX - 1) / TX;
  const int bx = lb % nbx, by = lb / nbx;
  const int i0 = bx * TX, j0 = by * TY;
  const int tid = threadIdx.x;

  // 1. keys
  for (int idx = tid; idx < KW * KH; idx += 256) {
    const int gi = i0 - 1 + idx % KW, gj = j0 - 1 + idx / KW;
    uint32_t k = 0;
    if (gi >= 0 && gj >= 0 && gi < g.vx && gj < g.vy) k = fkey(T.verts[(int64_t)gj * T.vpitch + gi]);
    sk[idx] = k;
  }
  __syncthreads();
  auto K = [&](int i, int j) -> uint64_t { return vkey(sk[(j - j0 + 1) * KW + (i - i0 + 1)], i, j); };

  // 2. candidates (DIR_NONE outside the grid or the interior rect)
  for (int idx = tid; idx < VW * VH; idx += 256) {
    const int i = i0 + idx % VW, j = j0 + idx / VW;
    int c = DIR_NONE;
    if (i < g.vx && j < g.vy && in_rect(g, 2 * i, 2 * j)) c = vertex_cand(K, g, i, j);
    svc[idx] = (uint8_t)c;
  }
  for (int idx = tid; idx < HW * HH; idx += 256) {
    const int i = i0 + idx % HW, j = j0 + idx / HW;
    int c = DIR_NONE;
    if (i + 1 < g.vx && j < g.vy && in_rect(g, 2 * i + 1, 2 * j)) c = hedge_cand(K, g, i, j);
    shc[idx] = (uint8_t)c;
  }
  for (int idx = tid; idx < EW * EH; idx += 256) {
    const int i = i0 + idx % EW, j = j0 + idx / EW;
    int c = DIR_NONE;
    if (i < g.vx && j + 1 < g.vy && in_rect(g, 2 * i, 2 * j + 1)) c = vedge_cand(K, g, i, j);
    sec[idx] = (uint8_t)c;
  }
  __syncthreads();
  auto VC = [&](int i, int j) -> int { return svc[(j - j0) * VW + (i - i0)]; };
  auto HC = [&](int i, int j) -> int { return shc[(j - j0) * HW + (i - i0)]; };
  auto EC = [&](int i, int j) -> int { return sec[(j - j0) * EW + (i - i0)]; };

  // 3. codes of the tile's cells
  for (int idx = tid; idx < CW * CH; idx += 256) {
    const int cx = 2 * i0 + idx % CW, cy = 2 * j0 + idx / CW;
    uint8_t c = 0;
    if (cx < g.cw && cy < g.ch) c = code_from_cands(K, VC, HC, EC, g, cx, cy);
    scode[idx] = c;
  }
  __syncthreads();

  // 4. block-local trace.  LDS word: 0x8000 | entry = link to another entry
  //    of this tile; otherwise the final owner value (packed root id,
  //    OWN_NONE, or OWN_MARK | id of the first cell outside the tile).
  for (int e = tid; e < 2 * NV; e += 256) {
    const bool quad = e >= NV;
    const int li = (e % NV) % TX, lj = (e % NV) / TX;
    const int lx = 2 * li + quad, ly = 2 * lj + quad;
    const int cx = 2 * i0 + lx, cy = 2 * j0 + ly;
    uint32_t w = OWN_NONE;
    if (cx < g.cw && cy < g.ch) {
      const uint8_t c = scode[ly * CW + lx];
      if (c & CF_CRITICAL) {
        w = pack_cell(T.ox + cx, T.oy + cy);
      } else if (c & CF_PAIRED) {
        const int d = code_dir(c);
        const int ni = li + dir_dx(d), nj = lj + dir_dy(d);
        if (ni >= 0 && nj >= 0 && ni < TX && nj < TY)
          w = 0x8000u | (uint32_t)((quad ? NV : 0) + nj * TX + ni);
        else
          w = OWN_MARK | pack_cell(T.ox + cx + 2 * dir_dx(d), T.oy + cy + 2 * dir_dy(d));
      }
    }
    sl[e] = w;
  }
  __syncthreads();
  for (;;) {
    int more = 0;
    for (int e = tid; e < 2 * NV; e += 256) {
      const uint32_t w = sl[e];
      if (lds_is_link(w)) {
        const uint32_t w2 = sl[w & 0x7FFFu];
        sl[e] = w2;
        more |= lds_is_link(w2);
      }
    }
    if (!__syncthreads_or(more)) break;
  }

  // 5. row stores: codes, owners, critical counts, marker list (the block's
  //    own region of T.exits, CW * CH entries, count in T.nexit[lb])
  if (tid == 0) s_nexit = 0;
  __syncthreads();
  uint32_t *const xlist = T.exits + (int64_t)lb * (CW * CH);
  const int lane = tid & 63, wv = tid >> 6;
  for (int r = wv; r < CH; r += 4) {
    const int cy = 2 * j0 + r;
    if (cy >= g.ch) break;
    const int lx = 2 * lane, cx = 2 * i0 + lx;
    const uint8_t c0 = scode[r * CW + lx], c1 = scode[r * CW + lx + 1];
    *(uint16_t *)(T.codes + (int64_t)cy * T.pitch + cx) = (uint16_t)(c0 | (c1 << 8));
    uint32_t o[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const uint8_t c = k ? c1 : c0;
      const int x = lx + k, gx = cx + k;
      uint32_t v = OWN_NONE;
      if (gx < g.cw) {
        const int dx = x & 1, dy = r & 1;
        if (dx == dy) {  // vertex (dx = 0) or quad (dx = 1)
          v = sl[(dx ? NV : 0) + (r >> 1) * TX + (x >> 1)];
        } else if (c & CF_CRITICAL) {
          v = pack_cell(T.ox + gx, T.oy + cy);
        } else if (c & CF_PAIRED) {
          const int d = code_dir(c);
          const int px = x + dir_dx(d), py = r + dir_dy(d);
          if (px >= 0 && py >= 0 && px < CW && py < CH)
            v = sl[((px & 1) ? NV : 0) + (py >> 1) * TX + (px >> 1)];
          else
            v = OWN_MARK | pack_cell(T.ox + gx + dir_dx(d), T.oy + cy + dir_dy(d));
        }
      }
      o[k] = v;
    }
    *(uint2 *)(T.own + (int64_t)cy * T.own_pitch + cx) = make_uint2(o[0], o[1]);
    const uint64_t crit_mask = __ballot((c0 >> 1) & 1) , crit_mask1 = __ballot((c1 >> 1) & 1);
    if (lane == 0) T.seg_counts[(int64_t)cy * T.nseg + bx] = (uint32_t)(__popcll(crit_mask) + __popcll(crit_mask1));
    const bool m0 = own_is_mark(o[0]), m1 = own_is_mark(o[1]);
    const uint64_t mm0 = __ballot(m0), mm1 = __ballot(m1);
    const uint32_t nm = (uint32_t)(__popcll(mm0) + __popcll(mm1));
    if (nm) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&s_nexit, nm);
      base = __shfl(base, 0, 64);
      const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
      uint32_t pos = base + (uint32_t)(__popcll(mm0 & lt) + __popcll(mm1 & lt));
      const uint32_t rowbase = (uint32_t)cy * (uint32_t)T.own_pitch + (uint32_t)cx;
      if (m0) xlist[pos++] = rowbase;
      if (m1) xlist[pos] = rowbase + 1;
    }
  }
  __syncthreads();
  if (tid == 0) T.nexit[lb] = s_nexit;
}

// Resolves the marker entries left by k_grad_trace: each follows the chain of
// markers (first cell outside a tile -> that cell's owner entry -> ...) to
// the owner at its end.  Chains follow gradient paths, so they end; the
// bound only guards against a corrupted image.
__global__ __launch_bounds__(256) void k_resolve_exits(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                       int ntiles, uint32_t *err) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc &T = tiles[t];
  const int lb = b - prefix[t];
  const uint32_t n = T.nexit[lb];
  const uint32_t *xlist = T.exits + (int64_t)lb * (4 * GRAD_TX * GRAD_TY);
  for (uint32_t e = threadIdx.x; e < n; e += 256) {
    const uint32_t idx = xlist[e];
    uint32_t w = T.own[idx];
    int steps = 0;
    while (own_is_mark(w) && steps < (1 << 24)) {
      const uint32_t c = w & ~OWN_MARK;
      const int x = (int)(c & 0xFFFF) - T.ox, y = (int)(c >> 16) - T.oy;
      w = T.own[(int64_t)y * T.own_pitch + x];
      ++steps;
    }
    if (own_is_mark(w)) atomicOr(err, 1u);
    T.own[idx] = w;
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
  const uint16_t word = *(const uint16_t *)(T.codes + (int64_t)row * T.pitch + seg * SEG + 2 * lane);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint64_t m0 = __ballot((word >> 1) & 1), m1 = __ballot((word >> 9) & 1);
  uint32_t idx = T.seg_offs[wid] + (uint32_t)(__popcll(m0 & lt) + __popcll(m1 & lt));
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const uint8_t c = (uint8_t)(word >> (8 * k));
    if (c & CF_CRITICAL) {
      const int cx = seg * SEG + 2 * lane + k, cy = row;
      if (idx < T.kcap) {
        T.crit[idx] = pack_cell(T.ox + cx, T.oy + cy);
        T.crit_fn[idx] = cell_fn(T, cx, cy);
      }
      ++idx;
    }
  }
}

// --------------------------------------------------------------- incidence
// Per critical cell: for saddles, the critical index owning each facet and
// each in-ext cofacet (writeout_connectivity, grid_dataset.cpp:1876-1898);
// for boundary-cancellable cells, the critical index of the pair.

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

__device__ __forceinline__ uint32_t own_at(const TileDesc &T, int cx, int cy) {
  return T.own[(int64_t)cy * T.own_pitch + cx];
}

__device__ __forceinline__ uint32_t crit_of_owner(const TileDesc &T, uint32_t K, uint32_t o) {
  return o == OWN_NONE ? 0xFFFFFFFFu : find_crit(T, K, (int)(o & 0xFFFF), (int)(o >> 16));
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
    // facets: the two vertices; cofacets: the quads across the edge in ext
    conn.x = crit_of_owner(T, K, own_at(T, cx + dx, cy + dy));
    conn.y = crit_of_owner(T, K, own_at(T, cx - dx, cy - dy));
    const int qx0 = cx + dy, qy0 = cy + dx, qx1 = cx - dy, qy1 = cy - dx;
    if (qx0 >= 0 && qy0 >= 0 && qx0 < g.cw && qy0 < g.ch) conn.z = crit_of_owner(T, K, own_at(T, qx0, qy0));
    if (qx1 >= 0 && qy1 >= 0 && qx1 < g.cw && qy1 < g.ch) conn.w = crit_of_owner(T, K, own_at(T, qx1, qy1));
  }
  ((uint4 *)T.crit_conn)[k] = conn;
  uint32_t pidx = ~0u;
  if (code & CF_PAIRED) {
    const int d = code_dir(code);
    pidx = find_crit(T, K, T.ox + cx + dir_dx(d), T.oy + cy + dir_dy(d));
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

void launch_grad_trace(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_grad_trace, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles);
}
void launch_resolve_exits(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, uint32_t *d_err,
                          hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_resolve_exits, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles, d_err);
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
void launch_incidence(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_incidence, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles);
}

}  // namespace mscx
