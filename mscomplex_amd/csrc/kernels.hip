// CDNA4 (gfx950) kernels for the per-tile discrete Morse-Smale stages:
// gradient assignment with a block-local manifold trace, resolution of the
// paths that cross block tiles, owner images, critical-cell collation (scan,
// compaction), saddle incidence and the manifold label pass.  All kernels
// take an array of tiles so one launch covers every tile of a tiling or of a
// batch of grids.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mscx {

// Device pointers taken from a tile descriptor are generic; marking them as
// global lets the compiler emit global (not flat) loads and stores.
#define GLB __attribute__((address_space(1)))
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
template <class T>
__device__ __forceinline__ GLB T *glb(T *p) {
  return (GLB T *)p;
}
template <class T>
__device__ __forceinline__ const GLB T *glb(const T *p) {
  return (const GLB T *)p;
}

// number of set bits of b in the lanes below this one, plus acc
__device__ __forceinline__ uint32_t rank_below(uint64_t b, uint32_t acc) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, acc));
}

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

__device__ __forceinline__ uint32_t pack_cell(int x, int y) { return (uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16); }

__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

// ------------------------------------------------- gradient + local trace
// One block = a GRAD_TX x TY vertex tile (cells 2*GRAD_TX x 2*TY; TY is the
// tile's block height, GRAD_TY or GRAD_TY_LOW rows);
// a lane owns one vertex column and, per vertex (i, j), the four cells
// (2i, 2j), (2i+1, 2j), (2i, 2j+1), (2i+1, 2j+1).
//  1. keys of the tile's vertices and a 1/2-vertex halo go to LDS (all
//     global loads of a lane are issued before the first LDS store; interior
//     blocks load without bounds tests)
//  2. each wave walks TY/4 rows with a 4x4 window of keys in registers
//     (rows j-1..j+2, columns i-1..i+2; one new LDS row per step) and
//     evaluates the pairing candidates it needs there (vertex_cand /
//     hedge_cand / vedge_cand, carrying the next row's over), then the four
//     cell codes (code_from_cands), the critical flags of each 128-cell row
//     segment (two ballots) and the trace links: a vertex points across its
//     paired edge to the next vertex, a quad across its paired edge to the
//     next quad (grid_dataset.cpp:1376-1427).  Blocks whose cells all lie
//     strictly inside the interior rect take a path without boundary tests,
//     with the walk unrolled.
//  3. pointer jumping in LDS takes every entry to its terminal in the tile:
//     its critical root, "none", or the exit through the tile's side
//  4. terminals go to `loc` (16 bits per vertex and per quad); the
//     terminals of the tile's perimeter entries go to its ring words, in the
//     form the ring resolution chases.

// The block's 16-bit trace entries in LDS: per forest (q = 0 vertices, 1
// quads) a (TX + 2) x (TY + 2) array whose interior holds entry (li, lj) at
// (li + 1, lj + 1) and whose one-entry border holds the exit terminals: the
// border slot a step from an edge entry lands on is T_EXIT with that side
// and position, so a link never needs to test whether it leaves the tile.
constexpr int ENT_W = GRAD_TX + 2;
template <int TY>
constexpr int ent_n() {
  return ENT_W * (TY + 2);
}
template <int TY>
__device__ __forceinline__ int ent_at(int q, int li, int lj) {
  return q * ent_n<TY>() + (lj + 1) * ENT_W + li + 1;
}

// Trace link of a vertex (q = 0) or quad (q = 1) entry (li, lj) from its
// code: when critical, its own root terminal (its cell row in the block,
// 2 lj + q, and its rank among the critical cells of that row's segment);
// none when unpaired; else the slot across its pair (DIR_PX, NX, PY, NY
// step by +1, -1, +ENT_W, -ENT_W), as a byte offset into the entry array (a
// border slot when the step leaves the tile).  Every value is computed
// before it is selected (no divergent branches).
template <int TY>
__device__ __forceinline__ uint32_t link_of(uint8_t c, int li, int lj, int q, uint32_t rank) {
  static_assert(ENT_W < 128, "step table holds +-ENT_W in a signed byte");
  constexpr uint32_t STEP = 0x01u | (0xFFu << 8) | ((uint32_t)ENT_W << 16) | ((uint32_t)(256 - ENT_W) << 24);
  const uint32_t root = T_ROOT | ((uint32_t)(2 * lj + q) << 7) | rank;
  const uint32_t s = (uint32_t)code_dir(c) - 1u;  // side 0..3 when paired
  const int step = __builtin_amdgcn_sbfe((int)STEP, (int)(8 * s), 8);
  const uint32_t link = 2u * (uint32_t)(ent_at<TY>(q, li, lj) + step);  // LDS byte offset
  const uint32_t paired = (c & CF_PAIRED) ? link : T_NONE;
  return (c & CF_CRITICAL) ? root : paired;
}

// link_of for an entry strictly inside the interior rect, from its pair
// direction d (0: critical, its root terminal).
template <int TY>
__device__ __forceinline__ uint32_t link_dir(int d, int li, int lj, int q, uint32_t rank) {
  constexpr uint32_t STEP = 0x01u | (0xFFu << 8) | ((uint32_t)ENT_W << 16) | ((uint32_t)(256 - ENT_W) << 24);
  const uint32_t root = T_ROOT | ((uint32_t)(2 * lj + q) << 7) | rank;
  const int step = __builtin_amdgcn_sbfe((int)STEP, 8 * (d - 1), 8);
  const uint32_t link = 2u * (uint32_t)(ent_at<TY>(q, li, lj) + step);
  return d ? link : root;
}

// Two interior cells' codes (interior_code of da | interior_code of db << 8)
// by one byte permute: selector byte k picks table byte k of the 8-byte
// table {2, 5, 9, 13 | 17, 0, 0, 0} (codes of directions 0..4).
__device__ __forceinline__ uint32_t interior_codes2(int da, int db) {
  static_assert(CF_PAIRED == 1 && CF_CRITICAL == 2 && DIR_NY == 4, "code table");
  constexpr uint32_t LO = 0x0D090502u, HI = 0x00000011u;
  return __builtin_amdgcn_perm(HI, LO, (uint32_t)da | ((uint32_t)db << 8)) & 0xFFFFu;
}

// Step 2 for one wave: rows lj0 .. lj0 + TY/4 - 1 of the block tile,
// with the row-step rules on a register window.  A lane computes the new
// candidates of its own column (row_own) and takes the next column's from
// the neighbouring lane; lane 63's next column lies outside the wave, so its
// candidates for all the wave's rows are computed once up front, one row per
// lane, and read with readlane.
template <int TY, bool INTERIOR>
__device__ __forceinline__ void grad_rows(const TileDesc &T, const TileGeom &g, const uint32_t *__restrict__ sk,
                                          uint16_t *__restrict__ sl, int i0, int j0, int bx, int lane, int lj0) {
  constexpr int TX = GRAD_TX, KW = TX + 3, RW = TY / 4;
  const int li = lane, i = i0 + li;
  // column i0 + TX: vertex candidate at row j0 + lj0 + L and vertical-edge
  // candidate at the same row, lane L in 0..RW (keys from LDS rows
  // lj0+L .. lj0+L+2, columns TX .. TX+2)
  int xv, xe;
  {
    const int L = lane < RW ? lane : RW;
    const uint32_t *s0 = sk + (lj0 + L) * KW + TX;
    KeyWin P;
    P.w00 = s0[0], P.w01 = s0[1], P.w02 = s0[2], P.w03 = 0;
    P.w10 = s0[KW], P.w11 = s0[KW + 1], P.w12 = s0[KW + 2], P.w13 = 0;
    P.w20 = s0[2 * KW], P.w21 = s0[2 * KW + 1], P.w22 = s0[2 * KW + 2], P.w23 = 0;
    P.w30 = P.w31 = P.w32 = P.w33 = 0;
    P.i = i0 + TX;
    P.j = j0 + lj0 + L;
    xv = vc_center<INTERIOR>(P, g);
    xe = ec_center<INTERIOR>(P, g);
  }
  KeyWin K;
  {
    const uint32_t *s0 = sk + lj0 * KW + li;
    K.w00 = s0[0], K.w01 = s0[1], K.w02 = s0[2], K.w03 = s0[3];
    K.w10 = s0[KW], K.w11 = s0[KW + 1], K.w12 = s0[KW + 2], K.w13 = s0[KW + 3];
    K.w20 = s0[2 * KW], K.w21 = s0[2 * KW + 1], K.w22 = s0[2 * KW + 2], K.w23 = s0[2 * KW + 3];
    K.w30 = s0[3 * KW], K.w31 = s0[3 * KW + 1], K.w32 = s0[3 * KW + 2], K.w33 = s0[3 * KW + 3];
  }
  K.i = i;
  K.j = j0 + lj0;
  int vc00 = vc_center<INTERIOR>(K, g);
  int hc00 = hc_center<INTERIOR>(K, g);
  int vc10 = __shfl_down(vc00, 1, 64);
  if (lane == 63) vc10 = __builtin_amdgcn_readlane(xv, 0);
  // this lane's two code bytes in cell row 2j, and this wave's critical
  // words / counts of cell row 2j; both advance by two cell rows per step
  const int64_t pitch = T.pitch, nseg = T.nseg;
  // (a tile's code array is below 4 GiB: a 32-bit offset from the uniform
  // base lets the stores address through a scalar base register)
  GLB uint8_t *const codes = glb(T.codes);
  uint32_t crow = (uint32_t)((int64_t)(2 * K.j) * pitch + 2 * i);
  int64_t rs0 = (int64_t)(2 * K.j) * nseg + bx;
  GLB uint64_t *const critbits = glb(T.critbits);
  // interior blocks lie inside the ext: every row exists, and the walk is
  // unrolled so that the window slides by register renaming
  constexpr int UNROLL = INTERIOR ? RW : 1;
#pragma unroll UNROLL
  for (int rr = 0; rr < RW; ++rr) {
    const int lj = lj0 + rr, j = K.j;
    if (!INTERIOR && j >= g.vy) break;  // uniform across the wave
    int vc01, hc01, ec00;
    row_own<INTERIOR>(K, g, vc01, hc01, ec00);
    // the next column's candidates from the next lane; lane 63's next
    // column is the prologue's column i0 + TX
    const int sv = __shfl_down(vc01, 1, 64), se = __shfl_down(ec00, 1, 64);
    const int xv1 = __builtin_amdgcn_readlane(xv, rr + 1), xe0 = __builtin_amdgcn_readlane(xe, rr);
    const int vc11 = lane == 63 ? xv1 : sv, ec10 = lane == 63 ? xe0 : se;
    if constexpr (INTERIOR) {
      // every cell is in the rect: a cell is critical exactly when it has
      // no pair, and its code is a function of its pair direction
      int dv, dh, de, dq;
      row_cell_dirs(vc00, vc10, hc00, vc01, hc01, ec00, ec10, dv, dh, de, dq);
      *(GLB uint16_t *)(codes + crow) = (uint16_t)interior_codes2(dv, dh);
      *(GLB uint16_t *)(codes + (crow + (uint32_t)pitch)) = (uint16_t)interior_codes2(de, dq);
      const uint64_t b0 = __ballot(dv == DIR_NONE), b1 = __ballot(dh == DIR_NONE);
      const uint64_t b2 = __ballot(de == DIR_NONE), b3 = __ballot(dq == DIR_NONE);
      if (lane == 0) {
        *(GLB u64x2 *)(critbits + 2 * rs0) = u64x2{b0, b1};
        *(GLB u64x2 *)(critbits + 2 * (rs0 + nseg)) = u64x2{b2, b3};
      }
      // ranks among the row segment's critical cells before this lane's
      // (the quad's also counts the edge of its own column: bits 0..lane of b2)
      const uint32_t rank_v = rank_below(b1, rank_below(b0, 0));
      const uint32_t rank_q = rank_below(b3, rank_below(b2 >> 1, (uint32_t)(b2 & 1u)));
      crow += 2 * (uint32_t)pitch;
      rs0 += 2 * nseg;
      sl[ent_at<TY>(0, li, lj)] = (uint16_t)link_dir<TY>(dv, li, lj, 0, rank_v);
      sl[ent_at<TY>(1, li, lj)] = (uint16_t)link_dir<TY>(dq, li, lj, 1, rank_q);
    } else {
      uint8_t cv, chh, ce, cq;
      row_cell_codes<INTERIOR>(K, g, vc00, vc10, hc00, vc01, hc01, ec00, ec10, cv, chh, ce, cq);
      const bool ein = j + 1 < g.vy;  // uniform across the wave
      *(GLB uint16_t *)(codes + crow) = (uint16_t)(cv | (chh << 8));
      const uint64_t b0 = __ballot(cv & CF_CRITICAL), b1 = __ballot(chh & CF_CRITICAL);
      // with no odd row below (ext edge) its words are not written
      const uint64_t b2 = __ballot(ein && (ce & CF_CRITICAL)), b3 = __ballot(ein && (cq & CF_CRITICAL));
      if (ein) *(GLB uint16_t *)(codes + (crow + (uint32_t)pitch)) = (uint16_t)(ce | (cq << 8));
      if (lane == 0) {
        *(GLB u64x2 *)(critbits + 2 * rs0) = u64x2{b0, b1};
        if (ein) *(GLB u64x2 *)(critbits + 2 * (rs0 + nseg)) = u64x2{b2, b3};
      }
      const uint32_t rank_v = rank_below(b1, rank_below(b0, 0));
      const uint32_t rank_q = rank_below(b3, rank_below(b2, 0)) + (uint32_t)((b2 >> lane) & 1u);
      crow += 2 * (uint32_t)pitch;
      rs0 += 2 * nseg;
      // codes outside the ext are 0, so those entries get T_NONE
      sl[ent_at<TY>(0, li, lj)] = (uint16_t)link_of<TY>(cv, li, lj, 0, rank_v);
      sl[ent_at<TY>(1, li, lj)] = (uint16_t)link_of<TY>(cq, li, lj, 1, rank_q);
    }
    vc00 = vc01;
    vc10 = vc11;
    hc00 = hc01;
    if (rr + 1 < RW) {
      const uint32_t *sn = sk + (lj + 4) * KW + li;
      K.slide(sn[0], sn[1], sn[2], sn[3]);
    } else {
      ++K.j;
    }
  }
}

// Keys of the block's vertices and their one-vertex halo into LDS: all of a
// lane's global loads are issued before its first LDS store.
template <int TY, bool INTERIOR>
__device__ __forceinline__ void load_keys(const TileDesc &T, const TileGeom &g, uint32_t *__restrict__ sk, int i0,
                                          int j0, int lane, int wv) {
  constexpr int TX = GRAD_TX, KW = TX + 3, KH = TY + 3;
  constexpr int NR = (KH + 3) / 4;  // rows per wave
  const GLB float *verts = glb((const float *)T.verts);
  const int64_t vpitch = T.vpitch;
  const int gi = i0 - 1 + lane, gi2 = i0 + TX - 1 + lane;
  float vm[NR], vh[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int r = wv + 4 * k;  // uniform across the wave
    const int gj = j0 - 1 + r;
    const bool rowok = r < KH && (INTERIOR || (gj >= 0 && gj < g.vy));
    const GLB float *row = verts + (int64_t)gj * vpitch;
    vm[k] = (rowok && (INTERIOR || (gi >= 0 && gi < g.vx))) ? row[gi] : 0.0f;
    vh[k] = (rowok && lane < 3 && (INTERIOR || gi2 < g.vx)) ? row[gi2] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int r = wv + 4 * k;
    if (r < KH) {
      sk[r * KW + lane] = fkey(vm[k]);
      if (lane < 3) sk[r * KW + TX + lane] = fkey(vh[k]);
    }
  }
}

// One block of the gradient kernel for a tile of TY-row block tiles; sk and
// sl are the launch's LDS arrays, sized for the tallest block.
template <int TY>
__device__ __forceinline__ void grad_block(const TileDesc &T, int lb, uint32_t *__restrict__ sk,
                                           uint16_t *__restrict__ sl, uint32_t *err) {
  constexpr int TX = GRAD_TX;
  constexpr int NV = TX * TY;  // entries per forest
  constexpr int RW = TY / 4;   // vertex rows per wave
  constexpr int ENT_N = ent_n<TY>();
  constexpr int RING_Q = ring_q(TY), RING_W = ring_w(TY);
  const TileGeom g = T.g;
  const int nbx = T.nseg;
  const int bx = lb % nbx, by = lb / nbx;
  const int i0 = bx * TX, j0 = by * TY;
  // the wave index is uniform: keep it (and the row indices derived from it)
  // in scalar registers
  const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);

  // A block is interior when every cell of it and of its one-vertex halo
  // lies strictly inside the interior rect: no boundary rule applies and
  // every key of the window is inside the ext.
  const bool cols_in = (2 * i0 - 1 > g.rx0) && (2 * (i0 + TX) + 1 < g.rx1);
  const bool interior = cols_in && (2 * j0 - 1 > g.ry0) && (2 * (j0 + TY) + 1 < g.ry1);
  // 1. keys (0 outside the ext window; the rules never read those)
  if (interior)
    load_keys<TY, true>(T, g, sk, i0, j0, lane, wv);
  else
    load_keys<TY, false>(T, g, sk, i0, j0, lane, wv);
  __syncthreads();

  // 2. codes, critical flags per row segment, trace links
  // (the row walk schedules better with its first row in a vector register)
  // A wave takes the interior walk when its own rows (with their one-vertex
  // halo, and the next wave's first row its walk also evaluates) are
  // interior, so boundary block rows run most of their waves without
  // boundary tests; the keys it reads are then all inside the ext.
  const int lj0 = wv * RW;
  const int wj0 = j0 + wv * RW;
  const bool wave_in = cols_in && (2 * wj0 - 1 > g.ry0) && (2 * (wj0 + RW) + 1 < g.ry1);
  if (wave_in)
    grad_rows<TY, true>(T, g, sk, sl, i0, j0, bx, lane, lj0);
  else
    grad_rows<TY, false>(T, g, sk, sl, i0, j0, bx, lane, lj0);
  // rows past the ext edge hold no entries
  for (int r = 0; r < RW; ++r) {
    const int lj = wv * RW + r;
    if (j0 + lj < g.vy) continue;
    sl[ent_at<TY>(0, lane, lj)] = (uint16_t)T_NONE;
    sl[ent_at<TY>(1, lane, lj)] = (uint16_t)T_NONE;
  }
  // the border: exit terminals of side s = 0 (+x), 1 (-x), 2 (+y), 3 (-y) at
  // the position along that side (the entry's row for x sides, column for
  // y).  Waves 0 / 1 fill the y sides of forest 0 / 1 (lane = column),
  // waves 2 / 3 the x sides (lane = row).
  {
    uint16_t *f = sl + (wv & 1) * ENT_N;
    const uint32_t pos = (uint32_t)lane;
    if (wv < 2) {
      f[(TY + 1) * ENT_W + lane + 1] = (uint16_t)(T_EXIT | (2u << 6) | pos);
      f[lane + 1] = (uint16_t)(T_EXIT | (3u << 6) | pos);
    } else if (lane < TY) {
      f[(lane + 1) * ENT_W + TX + 1] = (uint16_t)(T_EXIT | (0u << 6) | pos);
      f[(lane + 1) * ENT_W] = (uint16_t)(T_EXIT | (1u << 6) | pos);
    }
  }
  __syncthreads();

  // 3. pointer jumping.  Links only ever move further along their path, so
  //    the entries are updated without barriers: whatever value a read
  //    returns (the original link, a later one or the terminal) is a valid
  //    state.  A thread keeps its entries' current values in registers (it
  //    is the only writer of its entries); each round reads every target,
  //    then writes every entry, so the LDS latency is paid once per round.
  //    Terminal entries read a harmless in-range slot and keep their value.
  //    Gradient paths are acyclic and every round moves a pending entry at
  //    least one step, so the loop ends within the longest path (the bound
  //    guards bad input).
  {
    constexpr int PER = 2 * NV / 256;
    // links are byte offsets into sl, below the terminal bit; a terminal
    // masked by ENT_MASK is still an offset inside sl (T_ROOT's row and rank
    // bits, T_EXIT's 0x2000 with side and position, T_NONE's 0), so the read
    // for a terminal entry is harmless
    constexpr uint32_t ENT_MASK = 4 * ENT_N <= 0x2000 ? 0x1FFEu : 0x3FFEu;
    static_assert(4 * ENT_N <= ENT_MASK + 2 && ((((2u * TY - 1) << 7) | 127u) & ENT_MASK) < 4 * ENT_N &&
                      ((T_EXIT | 0xFFu) & ENT_MASK) < 4 * ENT_N,
                  "entry offsets and masked terminals");
    const char *slb = (const char *)sl;
    // this thread's entries: (li, lj) = (tid & 63, (tid >> 6) + 4 m) of
    // forest q, for k = MPER q + m; their slots are a fixed stride apart
    constexpr int MPER = TY / 4;
    const int e0 = ent_at<TY>(0, tid & 63, tid >> 6);
    auto slot = [&](int k) { return e0 + (k / MPER) * ENT_N + (k % MPER) * 4 * ENT_W; };
    static_assert(PER == 2 * MPER && TY % 4 == 0 && TX == 64, "entry walk layout");
    uint32_t cur[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) cur[k] = sl[slot(k)];
    uint32_t least = 0xFFFFu;  // all terminal: every entry >= 0x8000
#pragma unroll
    for (int k = 0; k < PER; ++k) least = min(least, cur[k]);
    for (int round = 0; least < 0x8000u && round < 2 * NV; ++round) {
      uint32_t nxt[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) nxt[k] = *(const uint16_t *)(slb + (cur[k] & ENT_MASK));
      least = 0xFFFFu;
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        cur[k] = cur[k] < 0x8000u ? nxt[k] : cur[k];
        sl[slot(k)] = (uint16_t)cur[k];
        least = min(least, cur[k]);
      }
    }
    const uint32_t all = least >= 0x8000u;
    if (!all) {
      atomicOr(err, 2u);
#pragma unroll
      for (int k = 0; k < PER; ++k)
        if (!(cur[k] & 0x8000u)) sl[slot(k)] = (uint16_t)T_NONE;
    }
  }
  __syncthreads();

  // 4. terminals and ring words
  for (int r = 0; r < RW; ++r) {
    const int lj = wv * RW + r;
    const int j = j0 + lj;
    if (j >= g.vy) break;
    T.loc[(int64_t)j * T.loc_pitch + i0 + lane] =
        (uint32_t)sl[ent_at<TY>(0, lane, lj)] | ((uint32_t)sl[ent_at<TY>(1, lane, lj)] << 16);
  }
  uint32_t *ring = T.ring + (int64_t)lb * RING_W;
  for (int p = tid; p < RING_W; p += 256) {
    const int q = p >= RING_Q ? 1 : 0;
    const int pp = p - q * RING_Q;
    int li, lj;
    ring_at(pp, TY, li, lj);
    uint32_t out = OWN_NONE;
    if (i0 + li + q < g.vx && j0 + lj + q < g.vy) {
      const uint32_t w = sl[ent_at<TY>(q, li, lj)];
      const uint32_t kind = term_kind(w);
      if (kind == 0) {
        // root: ext-local cell row, segment (= block column), rank in the segment
        const uint32_t y = (uint32_t)(2 * j0) + ((w >> 7) & 63u);
        out = (y << 15) | ((uint32_t)bx << 7) | (w & 127u);
      } else if (kind == 1) {
        const int s = (int)((w >> 6) & 3u), pos = (int)(w & 63u);
        const int tbx = bx + (s == 0) - (s == 1), tby = by + (s == 2) - (s == 3);
        // a path never leaves the ext window, so the neighbour exists
        if (tbx >= 0 && tby >= 0 && tbx < nbx && tby * TY < g.vy)
          out = RING_MARK | (uint32_t)((tby * nbx + tbx) * RING_W + q * RING_Q + ring_entry_base(s, TY) + pos);
        else
          atomicOr(err, 8u);
      }
    }
    ring[p] = out;
  }
}

__global__ __launch_bounds__(256) void k_grad_local(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                    int ntiles, uint32_t *err) {
  // keys of vertices [i0-1, i0+TX+1] x [j0-1, j0+TY+1] and the trace entries
  // of the tallest block tile
  __shared__ uint32_t sk[(GRAD_TY + 3) * (GRAD_TX + 3)];
  __shared__ uint16_t sl[2 * ent_n<GRAD_TY>()];
  static_assert(GRAD_TY_LOW < GRAD_TY, "tallest block");
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc T = tiles[t];
  if (T.ty == GRAD_TY_LOW)
    grad_block<GRAD_TY_LOW>(T, b - prefix[t], sk, sl, err);
  else
    grad_block<GRAD_TY>(T, b - prefix[t], sk, sl, err);
}

// Ring resolution: every perimeter entry follows ring words (one per block
// tile crossed) to its root and stores the root's critical index (its
// segment's scanned offset plus its rank there).
__global__ __launch_bounds__(256) void k_ring_resolve(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                      int ntiles, uint32_t *err) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc T = tiles[t];
  const int nblk = (int)tile_blocks(T.g, T.nseg, T.ty);
  const int64_t e = (int64_t)(b - prefix[t]) * 256 + threadIdx.x;
  if (e >= (int64_t)nblk * ring_w(T.ty)) return;
  uint32_t w = T.ring[e];
  for (int steps = 0; w != OWN_NONE && (w & RING_MARK); ++steps) {
    if (steps > nblk) {
      atomicOr(err, 4u);
      w = OWN_NONE;
      break;
    }
    w = T.ring[w & ~RING_MARK];
  }
  uint32_t out = OWN_NONE;
  if (w != OWN_NONE) out = T.seg_offs[(int64_t)(w >> 15) * T.nseg + ((w >> 7) & 255u)] + (w & 127u);
  T.ringidx[e] = out;
}

// Owner images: the terminals of the block's entries become critical indices
// (a root is its segment row's scanned offset plus its rank there, an exit
// the resolved ring word it leads to); an edge takes the owner of the vertex
// or quad it is paired with, a critical cell owns itself (its rank comes from
// ballots over the codes of its row).
template <int TY>
__device__ __forceinline__ void owner_block(const TileDesc &T, int lb, uint32_t *__restrict__ so,
                                            uint32_t *__restrict__ dv, uint32_t *__restrict__ dq,
                                            uint32_t *__restrict__ exit_base) {
  constexpr int TX = GRAD_TX, RW = TY / 4;
  constexpr int RING_Q = ring_q(TY), RING_W = ring_w(TY);
  const TileGeom g = T.g;
  const int nbx = T.nseg;
  const int bx = lb % nbx, by = lb / nbx;
  const int i0 = bx * TX, j0 = by * TY;
  // the wave index is uniform: keep it (and the row indices derived from it)
  // in scalar registers
  const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);

  const int64_t pitch = T.pitch, own_pitch = T.own_pitch, loc_pitch = T.loc_pitch;
  const GLB uint8_t *const codes = glb((const uint8_t *)T.codes);
  const GLB uint32_t *const ringidx = glb((const uint32_t *)T.ringidx);
  const GLB uint32_t *const loc = glb((const uint32_t *)T.loc);
  GLB uint32_t *const own = glb(T.own);
  uint32_t wl[RW];  // this lane's terminals (vertex | quad << 16) of its rows
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int j = j0 + wv * RW + r;
    wl[r] = (j < g.vy && i0 + lane < g.vx) ? loc[(int64_t)j * loc_pitch + i0 + lane] : (T_NONE | (T_NONE << 16));
  }
  for (int r = tid; r < 2 * TY; r += 256) {
    const int cy = 2 * j0 + r;
    if (cy < g.ch) so[r] = glb((const uint32_t *)T.seg_offs)[(int64_t)cy * nbx + bx];
  }
  // an exit through side s = 0 (+x), 1 (-x), 2 (+y), 3 (-y) at position p
  // continues at ring word exit_base[s] + q * RING_Q + p of the neighbour
  // (UINT32_MAX: no neighbour there; a valid gradient's paths stay inside
  // ext, so this only guards bad input)
  if (tid < 4) {
    const int tbx = bx + (tid == 0) - (tid == 1), tby = by + (tid == 2) - (tid == 3);
    const bool ok = tbx >= 0 && tby >= 0 && tbx < nbx && tby * TY < g.vy;
    exit_base[tid] = ok ? (uint32_t)(tby * nbx + tbx) * RING_W + (uint32_t)ring_entry_base(tid, TY) : 0xFFFFFFFFu;
  }
  __syncthreads();
  // ring word at offset off of forest q in block (tbx, tby), or UINT32_MAX
  // when that block does not exist (a valid gradient's paths stay inside
  // ext, so this only guards bad input)
  auto ring_index = [&](int tbx, int tby, int q, int off) -> uint32_t {
    if (tbx < 0 || tby < 0 || tbx >= nbx || tby * TY >= g.vy) return 0xFFFFFFFFu;
    return (uint32_t)(tby * nbx + tbx) * RING_W + (uint32_t)(q * RING_Q + off);
  };
  auto ring_val = [&](int tbx, int tby, int q, int off) -> uint32_t {
    const uint32_t x = ring_index(tbx, tby, q, off);
    return x == 0xFFFFFFFFu ? OWN_NONE : ringidx[x];
  };
  // 1. terminals -> critical indices in LDS.  A lane issues all its
  //    terminal loads, then all its ring loads, before it uses any of them.
  uint32_t rx[2 * RW];
#pragma unroll
  for (int k = 0; k < 2 * RW; ++k) {
    const int q = k & 1;
    const uint32_t w = (wl[k >> 1] >> (16 * q)) & 0xFFFFu;
    const uint32_t base = exit_base[(w >> 6) & 3u];
    const bool ok = term_kind(w) == 1 && base != 0xFFFFFFFFu;
    rx[k] = ok ? base + (uint32_t)(q * RING_Q) + (w & 63u) : 0xFFFFFFFFu;
  }
  uint32_t rv[2 * RW];
#pragma unroll
  for (int k = 0; k < 2 * RW; ++k) rv[k] = rx[k] != 0xFFFFFFFFu ? ringidx[rx[k]] : OWN_NONE;
  // owners across the block's sides that this lane's edges may be paired
  // with: per row, lane TX-1 the right block's vertex and lane 0 the left
  // block's quad; the block's first row the quads above it, its last row
  // the vertices below it
  uint32_t xs[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int lj = wv * RW + r;
    const uint32_t x = lane == TX - 1 ? ring_index(bx + 1, by, 0, ring_entry_base(0, TY) + lj)
                                      : (lane == 0 ? ring_index(bx - 1, by, 1, ring_entry_base(1, TY) + lj) : 0xFFFFFFFFu);
    xs[r] = x != 0xFFFFFFFFu ? ringidx[x] : OWN_NONE;
  }
  const uint32_t xt = wv == 0 ? ring_val(bx, by - 1, 1, ring_entry_base(3, TY) + lane) : OWN_NONE;
  const uint32_t xb = wv == 3 ? ring_val(bx, by + 1, 0, ring_entry_base(2, TY) + lane) : OWN_NONE;
#pragma unroll
  for (int k = 0; k < 2 * RW; ++k) {
    const int q = k & 1, lj = wv * RW + (k >> 1);
    const uint32_t w = (wl[k >> 1] >> (16 * q)) & 0xFFFFu;
    const uint32_t kind = term_kind(w);
    const uint32_t o = kind == 0 ? so[(w >> 7) & 63u] + (w & 127u) : kind == 1 ? rv[k] : OWN_NONE;
    (q ? dq : dv)[lj * TX + lane] = o;
  }
  // this lane's codes of all its rows, loaded while the others write LDS
  uint16_t cw0[RW], cw1[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int j = j0 + wv * RW + r;
    const GLB uint8_t *crow = codes + (int64_t)(2 * j) * pitch + 2 * (i0 + lane);
    cw0[r] = j < g.vy ? *(const GLB uint16_t *)crow : (uint16_t)0;
    cw1[r] = j + 1 < g.vy ? *(const GLB uint16_t *)(crow + pitch) : (uint16_t)0;
  }
  __syncthreads();

  // 2. owners of the four cells of each vertex: vertices and quads from
  //    the terminals, critical edges by rank, paired edges from the vertex or
  //    quad they are paired with (every value is computed, then selected).
  //    Unrolled, so that the per-row values held in register arrays (codes,
  //    side owners) are read without selects; rows past the ext edge are
  //    skipped by a branch that is uniform across the wave.
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int lj = wv * RW + r, li = lane;
    const int i = i0 + li, j = j0 + lj;
    if (j >= g.vy) continue;
    const bool ein = j + 1 < g.vy;
    const uint16_t w0 = cw0[r], w1 = cw1[r];
    const uint8_t cv = (uint8_t)w0, chh = (uint8_t)(w0 >> 8), ce = (uint8_t)w1, cq = (uint8_t)(w1 >> 8);
    const uint64_t bv = __ballot(cv & CF_CRITICAL), bh = __ballot(chh & CF_CRITICAL);
    const uint64_t be = __ballot(ce & CF_CRITICAL), bq = __ballot(cq & CF_CRITICAL);
    const uint32_t ov = dv[lj * TX + li], oq = dq[lj * TX + li];
    const uint32_t up_l = dq[(lj > 0 ? lj - 1 : 0) * TX + li], dn_l = dv[(lj + 1 < TY ? lj + 1 : lj) * TX + li];
    const uint32_t up_q = lj > 0 ? up_l : xt, dn_v = lj + 1 < TY ? dn_l : xb;
    const uint32_t rt_s = __shfl_down(ov, 1, 64), lf_s = __shfl_up(oq, 1, 64);
    const uint32_t rt_v = li == TX - 1 ? xs[r] : rt_s, lf_q = li == 0 ? xs[r] : lf_s;
    const uint32_t rh = so[2 * lj] + rank_below(bh, rank_below(bv, 0)) + (uint32_t)((bv >> lane) & 1u);
    const uint32_t re = so[2 * lj + 1] + rank_below(bq, rank_below(be, 0));
    const int dh = code_dir(chh), de = code_dir(ce);
    const uint32_t ph1 = dh == DIR_PY ? oq : up_q, ph2 = dh == DIR_NX ? ov : ph1, ph = dh == DIR_PX ? rt_v : ph2;
    const uint32_t pe1 = de == DIR_PX ? oq : lf_q, pe2 = de == DIR_NY ? ov : pe1, pe = de == DIR_PY ? dn_v : pe2;
    const uint32_t hp = (chh & CF_PAIRED) ? ph : OWN_NONE, ep = (ce & CF_PAIRED) ? pe : OWN_NONE;
    const uint32_t oh = (chh & CF_CRITICAL) ? rh : hp, oe = (ce & CF_CRITICAL) ? re : ep;
    GLB uint32_t *orow = own + (int64_t)(2 * j) * own_pitch + 2 * i;
    *(GLB u32x2 *)orow = u32x2{ov, oh};
    if (ein) *(GLB u32x2 *)(orow + own_pitch) = u32x2{oe, oq};
  }
}

__global__ __launch_bounds__(256) void k_owner(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                               int ntiles) {
  // LDS of the tallest block tile
  __shared__ uint32_t so[2 * GRAD_TY];
  __shared__ uint32_t dv[GRAD_TX * GRAD_TY], dq[GRAD_TX * GRAD_TY];
  __shared__ uint32_t exit_base[4];
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc T = tiles[t];
  if (T.ty == GRAD_TY_LOW)
    owner_block<GRAD_TY_LOW>(T, b - prefix[t], so, dv, dq, exit_base);
  else
    owner_block<GRAD_TY>(T, b - prefix[t], so, dv, dq, exit_base);
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

// Critical cells of (cell row, segment) item i: the popcount of its two
// critical words.
__device__ __forceinline__ uint32_t seg_count(const TileDesc &T, int64_t i) {
  const u64x2 w = ((const u64x2 *)T.critbits)[i];
  return (uint32_t)(__popcll(w[0]) + __popcll(w[1]));
}

__global__ __launch_bounds__(256) void k_scan_reduce(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                     int ntiles) {
  __shared__ uint32_t s_wave[4];
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc T = tiles[t];
  const int lb = b - prefix[t];
  const int64_t n = (int64_t)T.g.ch * T.nseg;
  const int64_t base = (int64_t)lb * SCAN_CHUNK;
  uint32_t s = 0;
  for (int k = 0; k < SCAN_CHUNK / 256; ++k) {
    int64_t i = base + k * 256 + threadIdx.x;
    if (i < n) s += seg_count(T, i);
  }
  uint32_t tot;
  block_exclusive_scan_256(s, s_wave, tot);
  if (threadIdx.x == 0) T.scan_part[lb] = tot;
}

__global__ __launch_bounds__(256) void k_scan_parts(const TileDesc *__restrict__ tiles) {
  __shared__ uint32_t s_wave[4];
  const TileDesc T = tiles[blockIdx.x];
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
  const TileDesc T = tiles[t];
  const int lb = b - prefix[t];
  const int64_t n = (int64_t)T.g.ch * T.nseg;
  const int64_t base = (int64_t)lb * SCAN_CHUNK + (int64_t)threadIdx.x * (SCAN_CHUNK / 256);
  uint32_t v[SCAN_CHUNK / 256];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_CHUNK / 256; ++k) {
    int64_t i = base + k;
    v[k] = (i < n) ? seg_count(T, i) : 0;
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
// One wave per 8 consecutive (cell row, segment) items; lane l holds cells
// 2l and 2l+1 of a segment: critical cells are written in raster order at
// their scanned offsets with their cell function value
// (grid_dataset.cpp:1538-1552: float sum in getCellPoints order, then the
// division).

constexpr int COMPACT_ITEMS = 8;  // (row, segment) items per wave

__global__ __launch_bounds__(256) void k_compact(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                 int ntiles) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc T = tiles[t];
  const int lane = threadIdx.x & 63;
  const int nseg = T.nseg, n = T.g.ch * nseg;
  const int first = __builtin_amdgcn_readfirstlane(((b - prefix[t]) * 4 + (int)(threadIdx.x >> 6)) * COMPACT_ITEMS);
  const GLB uint64_t *critbits = glb((const uint64_t *)T.critbits);
  const GLB uint32_t *seg_offs = glb((const uint32_t *)T.seg_offs);
  const GLB float *verts = glb(T.verts);
  const int64_t vpitch = T.vpitch;
  // 1. every item's flags and offset; then, for the items with critical
  //    cells, the corner values this lane's critical cells need (a = (i, j),
  //    b = (i+1, j), c = (i, j+1), d = (i+1, j+1)), all in flight together
  uint64_t fw[2 * COMPACT_ITEMS];
  uint32_t so[COMPACT_ITEMS];
#pragma unroll
  for (int it = 0; it < COMPACT_ITEMS; ++it) {
    const bool in = first + it < n;
    fw[2 * it] = in ? critbits[2 * (first + it)] : 0ull;
    fw[2 * it + 1] = in ? critbits[2 * (first + it) + 1] : 0ull;
    so[it] = in ? seg_offs[first + it] : 0u;
  }
  // (row, segment) of each item, advanced in scalar registers
  int srow[COMPACT_ITEMS], sseg[COMPACT_ITEMS];
  {
    int row = first / nseg, seg = first - row * nseg;
#pragma unroll
    for (int it = 0; it < COMPACT_ITEMS; ++it) {
      srow[it] = row, sseg[it] = seg;
      if (++seg == nseg) seg = 0, ++row;
    }
  }
  float va[COMPACT_ITEMS], vb[COMPACT_ITEMS], vc[COMPACT_ITEMS], vd[COMPACT_ITEMS];
#pragma unroll
  for (int it = 0; it < COMPACT_ITEMS; ++it) {
    va[it] = vb[it] = vc[it] = vd[it] = 0.0f;
    const uint64_t w0 = fw[2 * it], w1 = fw[2 * it + 1];
    if (!(w0 | w1)) continue;  // uniform across the wave
    const int seg = sseg[it], row = srow[it];
    const bool c0 = (w0 >> lane) & 1u, c1 = (w1 >> lane) & 1u;
    const int i = seg * (SEG / 2) + lane, j = row >> 1;
    const GLB float *p = verts + (int64_t)j * vpitch + i;
    if (c0 || c1) va[it] = p[0];
    if (c1) vb[it] = p[1];
    if ((row & 1) && (c0 || c1)) vc[it] = p[vpitch];
    if ((row & 1) && c1) vd[it] = p[vpitch + 1];
  }
  // 2. ranks, packed ids and cell values (get_cell_fn: float sum from 0 in
  //    getCellPoints order, then the division)
#pragma unroll
  for (int it = 0; it < COMPACT_ITEMS; ++it) {
    const uint64_t w0 = fw[2 * it], w1 = fw[2 * it + 1];
    if (!(w0 | w1)) continue;
    const int seg = sseg[it], row = srow[it];
    const uint32_t c0 = (uint32_t)(w0 >> lane) & 1u, c1 = (uint32_t)(w1 >> lane) & 1u;
    uint32_t k = rank_below(w1, rank_below(w0, so[it]));
    const int cx = seg * SEG + 2 * lane;
    const float a = va[it], bb = vb[it], c = vc[it], d = vd[it];
    float f0, f1;
    if (row & 1) {  // vertical edge (c, a), quad (d, b, a, c)
      f0 = ((0.0f + c) + a) / 2.0f;
      f1 = ((((0.0f + d) + bb) + a) + c) / 4.0f;
    } else {        // vertex (a), horizontal edge (b, a)
      f0 = (0.0f + a) / 1.0f;
      f1 = ((0.0f + bb) + a) / 2.0f;
    }
    if (c0 && k < T.kcap) {
      T.crit[k] = pack_cell(T.ox + cx, T.oy + row);
      T.crit_fn[k] = f0;
    }
    k += c0;
    if (c1 && k < T.kcap) {
      T.crit[k] = pack_cell(T.ox + cx + 1, T.oy + row);
      T.crit_fn[k] = f1;
    }
  }
}

// --------------------------------------------------------------- incidence
// Per critical cell: for saddles, the critical index owning each facet and
// each in-ext cofacet (writeout_connectivity, grid_dataset.cpp:1876-1898);
// for boundary-cancellable cells, the critical index of the pair (the pair
// is critical too, so its index is its rank in the scanned flags).

__device__ __forceinline__ uint32_t own_at(const TileDesc &T, int cx, int cy) {
  return T.own[(int64_t)cy * T.own_pitch + cx];
}

__device__ __forceinline__ uint32_t rank_at(const TileDesc &T, int cx, int cy) {
  const int64_t rs = (int64_t)cy * T.nseg + cx / SEG;
  return crit_rank(T.seg_offs[rs], T.critbits[2 * rs], T.critbits[2 * rs + 1], cx);
}

__global__ __launch_bounds__(256) void k_incidence(const TileDesc *__restrict__ tiles, const int *__restrict__ prefix,
                                                   int ntiles) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, ntiles, b);
  const TileDesc T = tiles[t];
  const TileGeom g = T.g;
  const uint32_t K = T.nkrit[0];
  const uint32_t k = (uint32_t)(b - prefix[t]) * 256 + threadIdx.x;
  if (k >= K || k >= T.kcap) return;
  const uint32_t c = T.crit[k];
  const int cx = (int)(c & 0xFFFF) - T.ox, cy = (int)(c >> 16) - T.oy;
  const uint8_t code = T.codes[(int64_t)cy * T.pitch + cx];
  const int dx = cx & 1, dy = cy & 1;
  uint4 conn = make_uint4(~0u, ~0u, ~0u, ~0u);
  if (dx != dy) {
    // facets: the two vertices; cofacets: the quads across the edge in ext
    conn.x = own_at(T, cx + dx, cy + dy);
    conn.y = own_at(T, cx - dx, cy - dy);
    const int qx0 = cx + dy, qy0 = cy + dx, qx1 = cx - dy, qy1 = cy - dx;
    if (qx0 >= 0 && qy0 >= 0 && qx0 < g.cw && qy0 < g.ch) conn.z = own_at(T, qx0, qy0);
    if (qx1 >= 0 && qy1 >= 0 && qx1 < g.cw && qy1 < g.ch) conn.w = own_at(T, qx1, qy1);
  }
  ((uint4 *)T.crit_conn)[k] = conn;
  uint32_t pidx = ~0u;
  if (code & CF_PAIRED) {
    const int d = code_dir(code);
    pidx = rank_at(T, cx + dir_dx(d), cy + dir_dy(d));
  }
  T.crit_pair[k] = pidx;
}

// ------------------------------------------------------------ disc labels
// For every interior vertex (quad) of a leaf: the label of its owner from
// the host's per-critical-index tables (a critical cell looks up its own
// entry, which applies its redirect first).

__global__ __launch_bounds__(256) void k_disc_labels(const DiscDesc *__restrict__ descs, const int *__restrict__ prefix,
                                                     int n) {
  // one block = 64 x 4 vertices of the interior rect; a thread labels its
  // vertex (2x, 2y) and the quad (2x+1, 2y+1) above and right of it
  const int b = blockIdx.x;
  const int t = find_tile(prefix, n, b);
  const DiscDesc D = descs[t];
  const int nvx = (D.rx1 - D.rx0) / 2 + 1, nvy = (D.ry1 - D.ry0) / 2 + 1;
  const int nbx = (nvx + 63) / 64, lb = b - prefix[t];
  const int x = (lb % nbx) * 64 + (int)(threadIdx.x & 63), y = (lb / nbx) * 4 + (int)(threadIdx.x >> 6);
  if (x >= nvx || y >= nvy) return;
  const bool hasq = x + 1 < nvx && y + 1 < nvy;
  const int lx = D.rx0 - D.ox + 2 * x, ly = D.ry0 - D.oy + 2 * y;
  const GLB uint32_t *own = glb(D.own);
  const GLB uint8_t *codes = glb(D.codes);
  const int64_t ov_at = (int64_t)ly * D.own_pitch + lx, cv_at = (int64_t)ly * D.pitch + lx;
  const uint32_t ov = own[ov_at], oq = hasq ? own[ov_at + D.own_pitch + 1] : OWN_NONE;
  const uint8_t cv = codes[cv_at], cq = hasq ? codes[cv_at + D.pitch + 1] : (uint8_t)0;
  auto label = [&](uint32_t o, uint8_t c) -> uint32_t {
    if (o == OWN_NONE) return 0xFFFFFFFFu;
    if (o >= D.nlab) {
      atomicOr(D.err, 2u);
      return 0xFFFFFFFFu;
    }
    const uint32_t lab = (c & CF_CRITICAL) ? D.lab_self[o] : D.lab_owned[o];
    if (lab == LAB_ERR) {
      atomicOr(D.err, 1u);
      return 0xFFFFFFFFu;
    }
    return lab;
  };
  const uint32_t lv = label(ov, cv), lq = label(oq, cq);
  D.lab_v[(int64_t)y * nvx + x] = lv;
  if (hasq) D.lab_q[(int64_t)y * (nvx - 1) + x] = lq;
}

// Sort input of the extremum discs' grouping (postMergeFillDiscs,
// grid_dataset.cpp:1142-1187, appends every vertex / quad of the rect to
// its extremum's disc in raster order): each vertex and quad of the rect,
// vertices then quads in raster order, keyed by its label plus the leaf's
// label base (unlabelled ones by the sentinel, which sorts last), valued by
// its cell id.  A stable sort by key then lists every disc's cells in
// raster order (one key per basin: counting them with atomics would
// serialise on a large basin's counter, so the offsets come from the sorted
// keys instead, k_disc_group_offsets).
__global__ __launch_bounds__(256) void k_disc_group_keys(const DiscDesc *__restrict__ descs,
                                                         const int *__restrict__ prefix, int n, uint32_t sentinel,
                                                         uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const int b = blockIdx.x;
  const int t = find_tile(prefix, n, b);
  const DiscDesc &D = descs[t];
  const int64_t nvx = (D.rx1 - D.rx0) / 2 + 1, nvy = (D.ry1 - D.ry0) / 2 + 1;
  const int64_t nv = nvx * nvy, nq = (nvx - 1) * (nvy - 1);
  const int64_t e = (int64_t)(b - prefix[t]) * 256 + threadIdx.x;
  if (e >= nv + nq) return;
  uint32_t lab;
  int x, y;
  if (e < nv) {
    lab = D.lab_v[e];
    x = D.rx0 + 2 * (int)(e % nvx);
    y = D.ry0 + 2 * (int)(e / nvx);
  } else {
    const int64_t e2 = e - nv, w = nvx - 1;
    lab = D.lab_q[e2];
    x = D.rx0 + 1 + 2 * (int)(e2 % w);
    y = D.ry0 + 1 + 2 * (int)(e2 / w);
  }
  const uint32_t key = lab == 0xFFFFFFFFu ? sentinel : D.lab_base + lab;
  keys[D.elem_base + e] = key;
  vals[D.elem_base + e] = pack_cell(x, y);
}

// Disc offsets: off[k] = the first position of a key >= k in the sorted
// keys, for k in [0, nkeys] (off[nkeys] = the labelled cells' count).
__global__ __launch_bounds__(256) void k_disc_group_offsets(const uint32_t *__restrict__ sorted, uint64_t n,
                                                            uint32_t nkeys, uint32_t *__restrict__ off) {
  const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (k > nkeys) return;
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (sorted[mid] < (uint32_t)k) lo = mid + 1;
    else hi = mid;
  }
  off[k] = (uint32_t)lo;
}

// ------------------------------------------------------- disc tracing
// postMergeFillDiscs' saddle 1-manifolds (grid_dataset.cpp:1189-1228).  The
// reference walks a growing list: every entry appends its successors (its
// cofacets inside the ext for the ascending disc, its facets for the
// descending one, that are not critical and lie in the rect, each replaced
// by its pair unless that is the entry itself).  Every entry after the
// initial ones is a paired cell whose own pair is one of those cofacets or
// facets, so it has at most one successor: the list is the initial entries,
// then the chains grown from their successors (in order) interleaved round
// by round, a chain dropping out when it ends.  A thread follows one slot's
// chains in that order, first counting, then writing.

__device__ __forceinline__ bool dt_in(int x, int y, int x0, int y0, int x1, int y1) {
  return x >= x0 && y >= y0 && x <= x1 && y <= y1;
}

// successors of a path cell; returns their number (at most 4)
__device__ __forceinline__ int dt_succ(const DiscTraceDesc &D, const GLB uint8_t *codes, uint32_t pc, int j,
                                       uint32_t succ[4]) {
  const int px = (int16_t)(pc & 0xFFFF), py = (int16_t)(pc >> 16);
  if (!dt_in(px, py, D.ox, D.oy, D.ex1, D.ey1)) return 0;
  const int dx = px & 1, dy = py & 1, dim = dx + dy;
  int cx[4], cy[4], ct = 0;
  if (j == 0) {
    if (dim == 0) {
      const int ox[4] = {0, 0, -1, 1}, oy[4] = {1, -1, 0, 0};
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (dt_in(px + ox[q], py + oy[q], D.ox, D.oy, D.ex1, D.ey1)) cx[ct] = px + ox[q], cy[ct] = py + oy[q], ++ct;
    } else if (dim == 1) {
      if (dt_in(px + dy, py + dx, D.ox, D.oy, D.ex1, D.ey1)) cx[ct] = px + dy, cy[ct] = py + dx, ++ct;
      if (dt_in(px - dy, py - dx, D.ox, D.oy, D.ex1, D.ey1)) cx[ct] = px - dy, cy[ct] = py - dx, ++ct;
    }
  } else {
    if (dim == 1) {
      cx[0] = px + dx, cy[0] = py + dy, cx[1] = px - dx, cy[1] = py - dy, ct = 2;
    } else if (dim == 2) {
      cx[0] = px, cy[0] = py + 1, cx[1] = px, cy[1] = py - 1, cx[2] = px - 1, cy[2] = py, cx[3] = px + 1, cy[3] = py;
      ct = 4;
    }
  }
  int ns = 0;
  for (int q = 0; q < ct; ++q) {
    const uint8_t code = codes[(int64_t)(cy[q] - D.oy) * D.pitch + (cx[q] - D.ox)];
    if (code & CF_CRITICAL) continue;
    if (!dt_in(cx[q], cy[q], D.rx0, D.ry0, D.rx1, D.ry1)) continue;
    const int d = code_dir(code);
    const uint32_t p = d ? pack_cell(cx[q] + dir_dx(d), cy[q] + dir_dy(d)) : 0xFFFFFFFFu;
    if (p == pc) continue;
    succ[ns++] = p;
  }
  return ns;
}

// One thread follows slot s's chains (count pass: its length, or the wide
// mark; write pass: its list).
template <bool WRITE>
__device__ void trace_slot(const DiscTraceDesc &D, uint32_t s) {
  const GLB uint8_t *codes = glb(D.codes);
  const uint32_t i0 = D.init_off[s], ni = D.init_off[s + 1] - i0;
  const uint32_t *init = D.init_cells + i0;
  // the write pass leaves the wide and host-traced slots to their tracers
  if (WRITE && D.count[s] >= DISC_TRACE_WIDE_LEN) return;
  GLB uint32_t *out = WRITE ? glb(D.out) + D.out_off[s] : nullptr;
  uint32_t len = 0;
  if (!D.traced[s] || ni == 0) {
    if (WRITE)
      for (uint32_t k = 0; k < ni; ++k) out[k] = init[k];
    else
      D.count[s] = ni;
    return;
  }
  const int j = (int)(s & 1u);
  uint32_t cur[DISC_TRACE_CHAINS];
  int r = 0;
  uint32_t succ[4];
  for (uint32_t k = 0; k < ni; ++k) {
    const uint32_t c = init[k];
    if (WRITE) out[len] = c;
    ++len;
    const int ns = dt_succ(D, codes, c, j, succ);
    for (int q = 0; q < ns; ++q) {
      if (r == DISC_TRACE_CHAINS) {
        // more chains than a thread follows: a wave traces this slot
        if (!WRITE) D.count[s] = DISC_TRACE_WIDE;
        return;
      }
      cur[r++] = succ[q];
    }
  }
  for (int k = 0; k < r; ++k)
    if (cur[k] == 0xFFFFFFFFu) {  // an unpaired successor: not a gradient the chains can follow
      atomicOr(D.err, 4u);
      return;
    }
  // round by round: the chains' current cells in order, each then replaced
  // by its successor (0xFFFFFFFF marks an ended chain).  A V-path of a
  // gradient is acyclic, so a chain is shorter than the ext has cells; the
  // bound ends the walk on invalid input.
  const uint32_t limit = (uint32_t)(D.ex1 - D.ox + 1) * (uint32_t)(D.ey1 - D.oy + 1);
  int alive = r;
  while (alive > 0) {
    for (int k = 0; k < r; ++k) {
      const uint32_t c = cur[k];
      if (c == 0xFFFFFFFFu) continue;
      if (WRITE) out[len] = c;
      ++len;
      const int ns = dt_succ(D, codes, c, j, succ);
      if (ns > 1 || (ns == 1 && succ[0] == 0xFFFFFFFFu) || len > limit + ni) {
        atomicOr(D.err, 2u);
        return;
      }
      cur[k] = ns ? succ[0] : 0xFFFFFFFFu;
      alive -= ns ? 0 : 1;
    }
  }
  if (!WRITE) D.count[s] = len;
}

// The wide slots (more chains than a thread follows): each wave takes the
// wide slots among its 64 one after another, with a lane per chain.  The list grows exactly as in
// k_disc_trace: the initial entries, then round by round every live chain's
// current cell in chain order; a round's positions are the ranks of the live
// chains (ballot and mbcnt over each group of 64 chains).  The chains' cells
// live in LDS (DISC_TRACE_WIDE_CAP per wave).
template <bool WRITE>
__device__ void trace_wide(const DiscTraceDesc &D, uint32_t s0, int lane, uint32_t *chains) {
  constexpr int CAP = DISC_TRACE_WIDE_CAP;
  const uint32_t my = s0 + (uint32_t)lane;
  const uint32_t mc = my < D.nslot ? D.count[my] : 0u;
  // count pass: the slots the thread pass marked; write pass: the slots the
  // wave's count pass measured
  const bool wide = WRITE ? (mc >= DISC_TRACE_WIDE_LEN && mc < DISC_TRACE_WIDE) : (mc == DISC_TRACE_WIDE);
  uint64_t todo = __ballot(wide);
  const GLB uint8_t *codes = glb(D.codes);
  const uint32_t limit = (uint32_t)(D.ex1 - D.ox + 1) * (uint32_t)(D.ey1 - D.oy + 1);
  while (todo) {
    // the chain table is the wave's own; the LDS accesses of one wave are
    // performed in order, so it is enough that the compiler keeps them so
    __builtin_amdgcn_wave_barrier();
    const int bit = __builtin_ctzll(todo);
    todo &= todo - 1;
    const uint32_t s = s0 + (uint32_t)bit;
    const uint32_t i0 = D.init_off[s], ni = D.init_off[s + 1] - i0;
    const uint32_t *init = D.init_cells + i0;
    GLB uint32_t *out = WRITE ? glb(D.out) + D.out_off[s] : nullptr;
    const int j = (int)(s & 1u);
    uint32_t succ[4];
    // the initial entries and their successors (the chains' heads, in the
    // order of their entries, then of the successors)
    uint32_t r = 0;
    bool over = false, bad_head = false;
    for (uint32_t k0 = 0; k0 < ni; k0 += 64) {
      const uint32_t k = k0 + (uint32_t)lane;
      const uint32_t c = k < ni ? init[k] : 0xFFFFFFFFu;
      if (WRITE && k < ni) out[k] = c;
      const int ns = k < ni ? dt_succ(D, codes, c, j, succ) : 0;
      // exclusive prefix of ns over the lanes
      int incl = ns;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
      }
      const uint32_t total = (uint32_t)__builtin_amdgcn_readlane(incl, 63);
      if (r + total > (uint32_t)CAP) {  // uniform
        over = true;
        break;
      }
      const uint32_t at = r + (uint32_t)(incl - ns);
      for (int q = 0; q < ns; ++q) {
        chains[at + q] = succ[q];
        bad_head |= succ[q] == 0xFFFFFFFFu;
      }
      r += total;
    }
    if (over) {
      if (!WRITE && lane == 0) D.count[s] = DISC_TRACE_HOST;
      continue;
    }
    __builtin_amdgcn_wave_barrier();
    if (__ballot(bad_head)) {  // an unpaired successor: not a gradient the chains can follow
      if (lane == 0) atomicOr(D.err, 4u);
      return;
    }
    // rounds: every live chain emits its cell and steps to its successor
    // (0xFFFFFFFF marks an ended chain)
    uint32_t len = ni;
    bool live = true, bad = false;
    while (live && !bad) {
      live = false;
      for (uint32_t g = 0; g < r; g += 64) {
        const uint32_t k = g + (uint32_t)lane;
        const uint32_t c = k < r ? chains[k] : 0xFFFFFFFFu;
        const bool alive = c != 0xFFFFFFFFu;
        const uint64_t bal = __ballot(alive);
        if (!bal) continue;
        live = true;
        if (WRITE && alive) out[rank_below(bal, len)] = c;
        len += (uint32_t)__popcll(bal);
        if (alive) {
          const int ns = dt_succ(D, codes, c, j, succ);
          bad |= ns > 1 || (ns == 1 && succ[0] == 0xFFFFFFFFu);
          chains[k] = ns ? succ[0] : 0xFFFFFFFFu;
        }
      }
      bad = __ballot(bad) != 0 || len > limit + ni;
    }
    if (bad) {
      if (lane == 0) atomicOr(D.err, 2u);
      return;
    }
    if (!WRITE && lane == 0) D.count[s] = DISC_TRACE_WIDE_LEN | len;
  }
}

// One launch per pass: a wave's lanes trace their own slots, then the wave
// traces the wide slots among them, so a pass lasts as long as its longest
// chain rather than the longest thread chain plus the longest wide one.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_disc_trace(const DiscTraceDesc *__restrict__ descs,
                                                    const int *__restrict__ prefix, int n) {
  __shared__ uint32_t chains_all[4][DISC_TRACE_WIDE_CAP];
  const int b = blockIdx.x;
  const int t = find_tile(prefix, n, b);
  const DiscTraceDesc &D = descs[t];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t s0 = (uint32_t)(b - prefix[t]) * 256u + (uint32_t)(wv * 64);
  if (s0 >= D.nslot) return;  // uniform across the wave
  if (s0 + (uint32_t)lane < D.nslot) trace_slot<WRITE>(D, s0 + (uint32_t)lane);
  // each lane reads back only the mark its own count pass wrote
  trace_wide<WRITE>(D, s0, lane, chains_all[wv]);
}

int plan_disc_trace_blocks(const DiscTraceDesc *d, int n, int *prefix) {
  int acc = 0;
  for (int i = 0; i < n; ++i) {
    prefix[i] = acc;
    acc += (int)((d[i].nslot + 255) / 256);
  }
  prefix[n] = acc;
  return acc;
}

void launch_disc_trace(const DiscTraceDesc *d_desc, const int *d_prefix, int n, int nblocks, bool write,
                       hipStream_t s) {
  if (!nblocks) return;
  if (write)
    hipLaunchKernelGGL(k_disc_trace<true>, dim3(nblocks), dim3(256), 0, s, d_desc, d_prefix, n);
  else
    hipLaunchKernelGGL(k_disc_trace<false>, dim3(nblocks), dim3(256), 0, s, d_desc, d_prefix, n);
}

// -------------------------------------------------------------- launchers

static int blocks_of(int kind, const TileDesc &T) {
  const TileGeom &g = T.g;
  const int64_t nblk = tile_blocks(g, T.nseg, T.ty);
  switch (kind) {
    case KB_GRADIENT:
      return (int)nblk;
    case KB_SCAN: {
      int64_t n = (int64_t)g.ch * T.nseg;
      return (int)((n + SCAN_CHUNK - 1) / SCAN_CHUNK);
    }
    case KB_RING:
      return (int)((nblk * ring_w(T.ty) + 255) / 256);
    case KB_COMPACT: {
      int64_t n = (int64_t)g.ch * T.nseg;
      return (int)((n + 4 * COMPACT_ITEMS - 1) / (4 * COMPACT_ITEMS));
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
    const int nvx = (d[t].rx1 - d[t].rx0) / 2 + 1, nvy = (d[t].ry1 - d[t].ry0) / 2 + 1;
    acc += ((nvx + 63) / 64) * ((nvy + 3) / 4);  // k_disc_labels: 64 x 4 vertices per block
  }
  prefix[n] = acc;
  return acc;
}

void launch_disc_labels(const DiscDesc *d_desc, const int *d_prefix, int n, int nblocks, hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_disc_labels, dim3(nblocks), dim3(256), 0, s, d_desc, d_prefix, n);
}
int plan_disc_group_blocks(const DiscDesc *d, int n, int *prefix) {
  int acc = 0;
  for (int t = 0; t < n; ++t) {
    prefix[t] = acc;
    const int64_t nvx = (d[t].rx1 - d[t].rx0) / 2 + 1, nvy = (d[t].ry1 - d[t].ry0) / 2 + 1;
    acc += (int)((nvx * nvy + (nvx - 1) * (nvy - 1) + 255) / 256);
  }
  prefix[n] = acc;
  return acc;
}
void launch_disc_group_keys(const DiscDesc *d_desc, const int *d_prefix, int n, int nblocks, uint32_t sentinel,
                            uint32_t *keys, uint32_t *vals, hipStream_t s) {
  if (nblocks)
    hipLaunchKernelGGL(k_disc_group_keys, dim3(nblocks), dim3(256), 0, s, d_desc, d_prefix, n, sentinel, keys, vals);
}
void launch_disc_group_offsets(const uint32_t *sorted, uint64_t n, uint32_t nkeys, uint32_t *off, hipStream_t s) {
  hipLaunchKernelGGL(k_disc_group_offsets, dim3((unsigned)((nkeys + 256) / 256)), dim3(256), 0, s, sorted, n, nkeys,
                     off);
}
void launch_grad_local(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, uint32_t *d_err,
                       hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_grad_local, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles, d_err);
}
void launch_ring_resolve(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, uint32_t *d_err,
                         hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_ring_resolve, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles, d_err);
}
void launch_owner(const TileDesc *d_tiles, const int *d_prefix, int ntiles, int nblocks, hipStream_t s) {
  if (nblocks) hipLaunchKernelGGL(k_owner, dim3(nblocks), dim3(256), 0, s, d_tiles, d_prefix, ntiles);
}
hipError_t build_front_graph(hipGraph_t *out, void *up_dst, const void *up_src, size_t up_bytes, void *dl_dst,
                             const void *dl_src, size_t dl_bytes, const TileDesc *d_tiles, const int *d_prefix,
                             int ntiles, const int *nblocks, uint32_t *d_err, const hipEvent_t *ev) {
  hipGraph_t g = nullptr;
  hipError_t e = hipGraphCreate(&g, 0);
  if (e != hipSuccess) return e;
  hipGraphNode_t prev = nullptr;
  auto chain = [&](hipGraphNode_t n) { prev = n; };
  auto memcpy_node = [&](void *dst, const void *src, size_t bytes, hipMemcpyKind kind) -> hipError_t {
    hipGraphNode_t n;
    hipError_t r = hipGraphAddMemcpyNode1D(&n, g, prev ? &prev : nullptr, prev ? 1 : 0, dst, src, bytes, kind);
    if (r == hipSuccess) chain(n);
    return r;
  };
  auto kernel_node = [&](const void *fn, int grid, void **args) -> hipError_t {
    if (grid <= 0) return hipSuccess;
    hipKernelNodeParams p{};
    p.func = const_cast<void *>(fn);
    p.gridDim = dim3(grid);
    p.blockDim = dim3(256);
    p.sharedMemBytes = 0;
    p.kernelParams = args;
    p.extra = nullptr;
    hipGraphNode_t n;
    hipError_t r = hipGraphAddKernelNode(&n, g, prev ? &prev : nullptr, prev ? 1 : 0, &p);
    if (r == hipSuccess) chain(n);
    return r;
  };
  auto event_node = [&](int k) -> hipError_t {
    if (!ev) return hipSuccess;
    hipGraphNode_t n;
    hipError_t r = hipGraphAddEventRecordNode(&n, g, prev ? &prev : nullptr, prev ? 1 : 0, ev[k]);
    if (r == hipSuccess) chain(n);
    return r;
  };
  const int *pg = d_prefix + KB_GRADIENT * (ntiles + 1), *ps = d_prefix + KB_SCAN * (ntiles + 1);
  const int *pr = d_prefix + KB_RING * (ntiles + 1);
  void *a_grad[] = {(void *)&d_tiles, (void *)&pg, (void *)&ntiles, (void *)&d_err};
  void *a_scan[] = {(void *)&d_tiles, (void *)&ps, (void *)&ntiles};
  void *a_parts[] = {(void *)&d_tiles};
  void *a_ring[] = {(void *)&d_tiles, (void *)&pr, (void *)&ntiles, (void *)&d_err};
  void *a_own[] = {(void *)&d_tiles, (void *)&pg, (void *)&ntiles};
  const bool scan = nblocks[KB_SCAN] > 0;
  if ((e = memcpy_node(up_dst, up_src, up_bytes, hipMemcpyHostToDevice)) == hipSuccess &&
      (e = event_node(0)) == hipSuccess &&
      (e = kernel_node((const void *)k_grad_local, nblocks[KB_GRADIENT], a_grad)) == hipSuccess &&
      (e = event_node(1)) == hipSuccess && (e = event_node(2)) == hipSuccess &&
      (e = kernel_node((const void *)k_scan_reduce, nblocks[KB_SCAN], a_scan)) == hipSuccess &&
      (e = kernel_node((const void *)k_scan_parts, scan ? ntiles : 0, a_parts)) == hipSuccess &&
      (e = kernel_node((const void *)k_scan_down, nblocks[KB_SCAN], a_scan)) == hipSuccess &&
      (e = event_node(3)) == hipSuccess && (e = event_node(4)) == hipSuccess &&
      (e = kernel_node((const void *)k_ring_resolve, nblocks[KB_RING], a_ring)) == hipSuccess &&
      (e = event_node(5)) == hipSuccess && (e = event_node(6)) == hipSuccess &&
      (e = kernel_node((const void *)k_owner, nblocks[KB_GRADIENT], a_own)) == hipSuccess &&
      (e = event_node(7)) == hipSuccess &&
      (e = memcpy_node(dl_dst, dl_src, dl_bytes, hipMemcpyDeviceToHost)) == hipSuccess) {
    *out = g;
    return hipSuccess;
  }
  (void)hipGraphDestroy(g);
  return e;
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
