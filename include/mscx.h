/* mscx.h -- C ABI of the MI355X Morse-Smale complex engine (libmscx.so).
 *
 * Drop-in boundary for the reference's GridDataset / GridMSComplex /
 * GridDataManager compute path (nithins/mscomplex, dmsc).  Every function
 * returns 0 on success and a negative status on failure; the message of the
 * last failure on the calling thread is returned by mscx_last_error().
 * Plain pointers and sizes only.  Cell ids are packed as x | (y << 16) with
 * x, y the int16 cell coordinates of the reference's cellid_t (the same bytes
 * as its two_tuple_t<int16_t> on little-endian hosts); (-1,-1) packs to
 * 0xFFFFFFFF.  Arrays over a tile are laid out over its ext rect, x fastest:
 * index = (x - ext[0]) + (y - ext[1]) * (ext[2] - ext[0] + 1).
 * Rects are int32_t[4] = {bl.x, bl.y, tr.x, tr.y} in cell coordinates
 * (vertex (i, j) is cell (2i, 2j)).
 */
#ifndef MSCX_H_INCLUDED
#define MSCX_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSCX_OK 0
#define MSCX_EINVAL -1    /* bad argument (reference: std::invalid_argument / range_error) */
#define MSCX_EDEVICE -2   /* HIP runtime failure (reference: runtime_error from _CHECKCL_ERR_CODE) */
#define MSCX_ELOGIC -3    /* invariant violated (reference: std::logic_error) */
#define MSCX_ENOMEM -4    /* allocation failure or HBM budget exceeded */
#define MSCX_ESTATE -5    /* call out of order (e.g. download before work) */
#define MSCX_EIO -6       /* input file cannot be opened or is too short (reference: runtime_error "unable to read stream") */
#define MSCX_ELIMIT -7    /* a served request reached its work limit (no reference counterpart) */

typedef struct mscx_tile mscx_tile;       /* a GridDataset: one rect + ext window on the GPU */
typedef struct mscx_complex mscx_complex; /* a GridMSComplex (host resident) */
typedef struct mscx_grid mscx_grid;       /* a GridDataManager run over one grid */

/* ---- runtime ------------------------------------------------------------
 * Replaces GridDataset::init_opencl / stop_opencl (grid_dataset.cpp:186-226,
 * grid_dataset.h:236-238).  mscx_init selects the HIP device; it is called
 * implicitly by the first tile call with device 0. */
int mscx_init(int device);
int mscx_shutdown(void);
const char *mscx_last_error(void);
int mscx_version(void);

/* Per-stage device timing (no reference counterpart; replaces the
 * reference's _LOG_TIMINGS wall-clock macros, grid_dataset.cpp:19-34).
 * mscx_profile(1) resets and enables HIP-event timing of each stage on the
 * work stream; mscx_profile_read returns accumulated milliseconds and event
 * pair counts for stages 0 gradient, 1 scan, 2 compaction, 3 trace,
 * 4 owners, 5 incidence, 6 host complex build, and the manifold pass's
 * parts as wall clock of the calls: 7 critical-point pass, 8 label tables,
 * 10 disc tracing on the GPU (with its transfers and scan), 13 grouping of
 * the extremum discs; and as host-thread time: 9 waiting for the labels,
 * 11 copying the traced lists, 12 grouping; 14 the wall clock of the
 * gradient codes' download that starts the manifold pass; and 15 the wall
 * clock of the leaves' host copies after a merge tree on the GPU. */
int mscx_profile(int enable);
int mscx_profile_read(double *ms, uint32_t *counts, int n);

/* Stage submission through HIP graphs (no reference counterpart).  With
 * on != 0 (the default) the part of a tile batch that needs no host decision
 * (descriptor upload, gradient, scan, ring resolution, owner images, count
 * download) is built into a HIP graph (node by node: no stream is put in
 * capture mode) once per launch configuration and replayed by later batches
 * with the same configuration; with 0 every stage
 * is launched directly.  While mscx_profile is enabled the graph also
 * records the stage timer's HIP events around its kernels, so the timed
 * submission path is the default one.  *previous (may be NULL) receives the
 * old setting;
 * mscx_graph_stats counts graph builds and replays since mscx_init. */
int mscx_set_graphs(int on, int *previous);
int mscx_graph_stats(uint64_t *builds, uint64_t *replays);

/* Block tile height of the gradient, trace and owner kernels (no reference
 * counterpart; the reference has no device kernels).  A tile whose window
 * has vy vertex rows takes rows[k] of the first class k with
 * vy <= max_rows[k] (at most 16 classes, bounds increasing) and 32 rows when
 * no class matches; rows[k] is 32 or 16.  Blocks are 64 vertex columns wide
 * (one per lane of a wavefront).  Tiles created afterwards use the table;
 * n = 0 restores the default.  Results do not depend on it.  The batched
 * form's autotuner (mscomplex_amd.autotune_block_rows) measures both
 * heights per class and installs its choice here. */
int mscx_set_block_rows(int32_t n, const int32_t *max_rows, const int32_t *rows);

/* Counters of the manifold pass's disc tracing since the library loaded (no
 * reference counterpart): [0] disc slots traced on the GPU, [1] slots
 * traced on the host (a disc whose initial entries grow more chains than a
 * wavefront's chain table holds), [2] calls whose GPU lists were dropped for
 * the host's because the chains were not a gradient's V-paths (never
 * expected), [3] GPU slots traced by a whole wavefront (more chains than
 * one lane follows), [4] the longest list traced on the GPU (cells; a
 * tracing pass lasts about as long as its chain of dependent reads). */
int mscx_disc_trace_stats(uint64_t out[5]);
/* Counters of the device merge tree since the library loaded (no reference
 * counterpart): merges run on the GPU (up and down), grids whose tree went
 * back to the host path (a case the kernels do not follow), and host
 * passes over a shared line. */
int mscx_merge_stats(uint64_t out[3]);

/* Device buffers (no reference counterpart: the reference keeps inputs in
 * host memory).  Lets a caller keep grids resident in HBM and bind tiles to
 * them with mscx_tile_bind_device. */
int mscx_dev_alloc(size_t bytes, void **out);
int mscx_dev_free(void *p);
int mscx_dev_upload(void *dst, const void *src, size_t bytes);
int mscx_dev_download(void *dst, const void *src, size_t bytes);

/* ---- tiles: GridDataset ------------------------------------------------- */

/* GridDataset(const rect_t &r, const rect_t &e) (grid_dataset.h:97,
 * grid_dataset.cpp:1429-1441).  ext must contain rect; both have even
 * corners; at most 16384 vertices per side (int16 cell ids). */
int mscx_tile_create(const int32_t rect[4], const int32_t ext[4], mscx_tile **out);
int mscx_tile_destroy(mscx_tile *t);

/* GridDataset::init(float *) (grid_dataset.cpp:1461-1484): the ext window
 * of vertex values, x fastest, (ext_w/2+1) * (ext_h/2+1) floats, copied to
 * HBM.  NaN values are rejected with MSCX_EINVAL. */
int mscx_tile_upload(mscx_tile *t, const float *x_fastest, size_t n);

/* Device-resident variant of init(): dev_ptr points to the window's first
 * vertex in HBM with row_pitch floats per row.  No copy, no NaN check. */
int mscx_tile_bind_device(mscx_tile *t, const float *dev_ptr, int64_t row_pitch);

/* GridDataset::work() / work_ocl() (grid_dataset.cpp:1734-1741, :353-399):
 * gradient, critical collation, owner tracing and saddle incidence on the
 * GPU, synchronously. */
int mscx_tile_work(mscx_tile *t);

/* Batched work(): one launch per stage over all n tiles. */
int mscx_tiles_work(mscx_tile *const *tiles, int n);

/* mscx_tile_work without any device allocation.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It runs the same kernels as mscx_tile_work on one tile
 * over (rect, ext) whose input window is already in HBM (dev_verts, row_pitch
 * floats per row) and writes the tile's critical cells into the caller's
 * device arrays of kcap entries: dev_crit (packed ids, raster order),
 * dev_fn (get_cell_fn values), dev_conn4 (4 per cell, the layout of
 * mscx_tile_download_critical) and dev_pair; *K (host) receives K.  The
 * tile's intermediates live in a static 32 MiB device array per GPU, so the
 * tile may have at most about 1026 x 1026 vertices (larger tiles fail with
 * MSCX_EINVAL); K > kcap fails with MSCX_EINVAL after writing K to *K.  The
 * call allocates no device or pinned host memory; calls on one GPU share its
 * array and are serialised by the library (the array keeps the last call's
 * contents, and callers that bypass the library's lock on one GPU race). */
int mscx_tile_work_noalloc(const int32_t rect[4], const int32_t ext[4], const float *dev_verts, int64_t row_pitch,
                           uint32_t kcap, uint32_t *dev_crit, float *dev_fn, uint32_t *dev_conn4,
                           uint32_t *dev_pair, uint32_t *K);

/* Number of critical cells K (m_critical_cells.size()). */
int mscx_tile_num_critical(const mscx_tile *t, uint32_t *K);

/* Host copies of m_cell_flags (int8), m_cell_pairs and m_cell_own (packed
 * ids) over the ext rect, and m_critical_cells (K packed ids, raster order).
 * Any pointer may be NULL. */
int mscx_tile_download(const mscx_tile *t, int8_t *flags, uint32_t *pairs, uint32_t *own, uint32_t *crit);

/* Raw per-cell gradient codes (bits 0-1 flags, bits 2-4 pair direction
 * 0 none, 1 +x, 2 -x, 3 +y, 4 -y) over the ext rect, x fastest. */
int mscx_tile_download_codes(const mscx_tile *t, uint8_t *codes);

/* Per critical cell: function value (get_cell_fn), incident owners of a
 * saddle (two facet owners then two cofacet owners, as critical indices or
 * 0xFFFFFFFF), and the critical index of the pair of a boundary-cancellable
 * cell (or 0xFFFFFFFF).  Arrays of K, K, 4K, K entries; any may be NULL. */
int mscx_tile_download_critical(const mscx_tile *t, float *fn, uint32_t *conn4, uint32_t *pair_idx);

/* ---- complexes: GridMSComplex -------------------------------------------- */

/* GridDataset::writeout_connectivity (grid_dataset.cpp:1858-1907): builds a
 * complex from a tile after mscx_tile_work. */
int mscx_complex_from_tile(const mscx_tile *t, mscx_complex **out);

/* GridDataset::writeout_connectivity_ocl counterpart (grid_dataset.cpp:
 * 1250-1283): builds a complex from per-critical-cell data in the layout of
 * mscx_tile_download_critical (K packed ids in raster order, K values, 4K
 * incident owner indices, K pair indices).  Host only. */
int mscx_complex_from_critical(const int32_t rect[4], const int32_t ext[4], uint32_t K, const uint32_t *crit,
                               const float *fn, const uint32_t *conn4, const uint32_t *pair_idx,
                               mscx_complex **out);

/* GridMSComplex::merge_up (grid_mscomplex.cpp:24-156); the caller owns *out. */
int mscx_complex_merge_up(const mscx_complex *a, const mscx_complex *b, mscx_complex **out);

/* GridMSComplex::merge_down (grid_mscomplex.cpp:158-322). */
int mscx_complex_merge_down(mscx_complex *parent, mscx_complex *a, mscx_complex *b);

/* GridMSComplex::simplify_un_simplify (grid_mscomplex.cpp:528-535). */
int mscx_complex_simplify_un_simplify(mscx_complex *c, double threshold);

/* GridDataset::postMergeFillDiscs (grid_dataset.cpp:1044-1231): fills the
 * manifold discs of the tile's complex from the tile's owner image. */
int mscx_tile_fill_discs(mscx_tile *t, mscx_complex *c);

/* postMergeFillDiscs from host copies of a tile's gradient codes and owner
 * image over its ext rect (x fastest); own is modified in place. */
int mscx_complex_fill_discs_host(mscx_complex *c, const int32_t rect[4], const int32_t ext[4],
                                 const uint8_t *codes, uint32_t *own);

/* GridMSComplex::write_discs (grid_mscomplex.cpp:542-590): appends raw
 * packed cell ids to <prefix>asc_(x,y) / <prefix>des_(x,y). */
int mscx_complex_write_discs(const mscx_complex *c, const char *prefix);

/* Canonical dump of a complex:
 *   [ncp, rect[4], ext[4]] then per critical point: cellid, fn bits,
 *   pair_idx, flags (1 cancelled, 2 strangulation, 4 boundary-cancellable),
 *   n_asc, n_des, n_asc_disc, n_des_disc, asc..., des..., asc_disc...,
 *   des_disc...
 * *len receives the word count; the words are written when cap suffices. */
int mscx_complex_dump(const mscx_complex *c, uint32_t *buf, size_t cap, size_t *len);
/* Host archive of a complex: the counterpart of write_msgraph_to_archive /
 * read_msgraph_from_archive (grid_datamanager.cpp:91-119), which store a
 * GridMSComplex's rect, ext, values, id map and critical points
 * (grid_mscomplex.cpp:604-644) as Boost binary archives.  The format is
 * this library's own and portable across little-endian hosts: magic
 * "MSCXCPX1", version 1, a reserved word, the word count (uint64), the dump
 * words above (discs included) and an FNV-1a 64 checksum of them.  A loaded
 * complex dumps word for word as the saved one and takes part in merges,
 * simplification and disc writing like any other.  Load fails with
 * MSCX_EIO on a missing, truncated, corrupt or foreign file. */
int mscx_complex_save(const mscx_complex *c, const char *path);
int mscx_complex_load(const char *path, mscx_complex **out);
int mscx_complex_num_cps(const mscx_complex *c, uint32_t *n);
int mscx_complex_destroy(mscx_complex *c);

/* ---- whole-grid pipeline: GridDataManager -------------------------------- */

/* GridDataManager(filename, size_x, size_y, num_levels, ..., simp_tresh, ...)
 * compute order (grid_datamanager.cpp:596-662): tile the grid, compute every
 * leaf on the GPU in one batched launch per stage, merge up, simplify the
 * root at `threshold` (fraction of the critical-value range; 0 = none),
 * merge down, and fill manifold discs.  levels = slab levels along y as in
 * createPieces_quadtree (grid_datamanager.cpp:53-89); lx/ly > 0 selects a
 * 2^lx x 2^ly 2-D tiling.  data is host (x fastest) unless
 * MSCX_RUN_DEVICE_INPUT is set, then it is a device pointer. */
#define MSCX_RUN_DEVICE_INPUT 1
#define MSCX_RUN_NO_DISCS 2
/* The reference's --out-of-core (grid_datamanager.cpp:91-119,134-210): a
 * piece that no phase is using is kept serialised -- as its archive words in
 * host memory, or as an archive file under the directory named by the
 * MSCX_SPILL_DIR environment variable -- and brought back when a phase or
 * mscx_grid_piece needs it, so the merge tree's host memory is bounded by
 * the pieces one merge level works on rather than by every level at once
 * (a 16384^2 white-noise grid's tree holds ~228 M critical points per
 * level).  Results are the same as without the flag. */
#define MSCX_RUN_OUT_OF_CORE 4
/* The merge tree on the GPU: the pieces' complexes stay in HBM, merge_up
 * and merge_down (grid_mscomplex.cpp:24-322) run as kernels (the shared
 * line's cancellations on the host over the line's gathered sub-complex),
 * the root comes to the host for the sequential simplification, and a
 * piece is copied to the host when a phase or mscx_grid_piece needs it.
 * Results are the same as without the flag (a merge the kernels do not
 * follow sends the whole tree to the host path).  mscx_grid_run and
 * mscx_grid_run_file set it by themselves for grids of at least
 * MSCX_DEVICE_MERGE_MIN_VERTICES vertices (environment; default 2^25:
 * smaller trees are latency-bound on the GPU, where the host tree is as
 * fast);
 * MSCX_RUN_HOST_MERGE keeps the tree on the host. */
#define MSCX_RUN_DEVICE_MERGE 8
#define MSCX_RUN_HOST_MERGE 16
int mscx_grid_run(const float *data, int32_t size_x, int32_t size_y, int32_t levels, int32_t lx, int32_t ly,
                  double threshold, uint32_t flags, mscx_grid **out);
/* The whole reference driver over a raw grid file: GridDataManager(filename,
 * size_x, size_y, num_levels, single_threaded, use_ocl, simp_tresh,
 * out_of_core, num_parallel) (grid_datamanager.cpp:596-662) with its reader
 * readDataAndInit / computeSubdomainMsgraphs (:493-594) and, when
 * disc_prefix is not NULL, collectManifold's disc files
 * (<disc_prefix>asc_(x,y) / des_(x,y), appended leaf by leaf in piece order;
 * the reference driver's prefix is "dp__disc_", :397-420).  The file holds
 * at least size_x * size_y little-endian floats, x fastest (the reference's
 * format, main.cpp -gf); trailing bytes are ignored as the reference's
 * reader ignores them.  It is streamed once into HBM in row slabs through
 * two pinned host buffers, each slab's copy overlapping the next read
 * (MSCX_READ_SLAB_BYTES in the environment sets the slab size, default
 * 8 MiB); the pieces bind to the resident grid, so the reference's
 * three-row rewind between buffers has no counterpart.  num_parallel caps
 * the host threads of the host phases (0 = default, 1 = the reference's
 * -st).  levels / lx / ly / threshold / flags as for mscx_grid_run
 * (MSCX_RUN_DEVICE_INPUT is invalid here).  A missing, unreadable or short
 * file fails with MSCX_EIO before the device is touched. */
int mscx_grid_run_file(const char *path, int32_t size_x, int32_t size_y, int32_t levels, int32_t lx, int32_t ly,
                       double threshold, int32_t num_parallel, uint32_t flags, const char *disc_prefix,
                       mscx_grid **out);
int mscx_grid_num_pieces(const mscx_grid *g, int32_t *n);
/* Borrowed pointer to piece i's complex (leaves first, root last). */
int mscx_grid_piece(const mscx_grid *g, int32_t i, const mscx_complex **out);
/* Number of critical points of piece i without making a host copy of a
 * device-resident piece (no reference counterpart). */
int mscx_grid_piece_num_cps(const mscx_grid *g, int32_t i, uint32_t *n);
/* FNV-1a checksum of the dump words of every piece in piece order (no
 * reference counterpart): equal checksums of two runs mean their pieces
 * agree word for word, for comparing forms (batched, served, plain) at
 * sizes where keeping every dump is too costly. */
int mscx_grid_checksum(const mscx_grid *g, uint64_t *out);
/* Stage times in seconds: compute, merge-up, simplify, merge-down, discs. */
int mscx_grid_times(const mscx_grid *g, double t5[5]);
int mscx_grid_destroy(mscx_grid *g);

/* Many independent grids at once (no reference counterpart: the reference
 * handles one grid per GridDataManager).  Each stage's GPU work for every
 * leaf tile of every grid is one launch; the host phases of different grids
 * run concurrently.  Per-grid arrays of length n: data (host, or device with
 * MSCX_RUN_DEVICE_INPUT), sizes, 2-D tiling levels and thresholds.  A grid
 * that fails (bad input, arc explosion) gets its own status; the others
 * complete.  mscx_batch_grid returns a borrowed mscx_grid for the grid
 * accessors above. */
typedef struct mscx_batch mscx_batch;
int mscx_batch_run(int32_t n, const float *const *data, const int32_t *size_x, const int32_t *size_y,
                   const int32_t *lx, const int32_t *ly, const double *threshold, uint32_t flags, mscx_batch **out);
int mscx_batch_size(const mscx_batch *b, int32_t *n);
int mscx_batch_grid(const mscx_batch *b, int32_t i, const mscx_grid **out, int32_t *status);
const char *mscx_batch_error(const mscx_batch *b, int32_t i);
int mscx_batch_destroy(mscx_batch *b);

/* ---- serving: continuous batching over a request queue --------------------
 * No reference counterpart (the reference runs one grid per GridDataManager);
 * every result equals mscx_grid_run on the same grid, 2-D tiling and
 * threshold.  A scheduler thread admits queued requests against an HBM
 * budget (bytes of device memory the library may hold) and launches the GPU
 * stages of all grids admitted in one round together; their host phases run
 * on a pool of host_threads workers (0 = one per core, at most 16) while the
 * next round is admitted, so requests that arrive while others are in
 * flight overlap with them.  A request with a nonzero key leaves its grid
 * resident (owner images in HBM, merged complexes on the host) until the
 * budget needs the memory (least recently used first); a later request with
 * the same key is served from it -- simplified at its own threshold without
 * the GPU stages -- and may pass data = NULL while the key is resident.  The
 * key names the grid and its tiling.  A grid larger than the budget fails
 * with MSCX_ENOMEM. */
typedef struct mscx_server mscx_server;
int mscx_server_create(size_t hbm_budget, int32_t host_threads, mscx_server **out);
/* Work limit of the requests' simplification (no reference counterpart:
 * the reference runs until done, which on noisy fields at large thresholds
 * takes minutes to hours as strangulated pairs are resubmitted after every
 * cancellation, and grows arc multisets without bound).  A request whose
 * simplification needs more than max(2^24, factor x its initial heap
 * entries) heap pops (at most 2^27, a few seconds) fails with MSCX_ELIMIT, one
 * that would insert more than
 * max(2^24, 2 x factor x its arc entries) arc entries with MSCX_ENOMEM;
 * both in seconds, so the worker moves on.  Default 128 (the 4096^2 value
 * noise root at 1% takes 35x); 0 removes the limit.  Applies to requests
 * that start after the call.  Plain and batched runs take the same limit
 * from the MSCX_WORK_LIMIT environment variable (unset or 0: none). */
int mscx_server_set_work_limit(mscx_server *s, double factor);
int mscx_server_submit(mscx_server *s, const float *data, int32_t size_x, int32_t size_y, int32_t lx, int32_t ly,
                       double threshold, uint64_t key, uint32_t flags, uint64_t *ticket);
/* Waits up to timeout_s seconds (< 0: no limit) for a request; *out is
 * borrowed until mscx_server_release, *status its own status, *latency_s
 * the time from submission to completion. */
int mscx_server_wait(mscx_server *s, uint64_t ticket, double timeout_s, const mscx_grid **out, int32_t *status,
                     double *latency_s);
const char *mscx_server_error(mscx_server *s, uint64_t ticket);
int mscx_server_release(mscx_server *s, uint64_t ticket);
/* done, scheduling rounds, repeat requests, evictions, resident grids,
 * device bytes held, budget, queued requests */
int mscx_server_stats(mscx_server *s, uint64_t out[8]);
int mscx_server_destroy(mscx_server *s);

/* ---- synthetic inputs (seeded, host) -------------------------------------- */

/* Sum of n Gaussians, centres U[0,sx)xU[0,sy), sigma U[8,64], amplitude
 * N(0,1), plus U[-noise, noise]. */
int mscx_gen_gaussians(float *out, int32_t sx, int32_t sy, int32_t n, double noise, uint64_t seed);
/* Fractal value noise: octaves, lacunarity, gain, lattice values U[0,1),
 * base period `period` pixels, smoothstep-bilinear interpolation. */
int mscx_gen_value_noise(float *out, int32_t sx, int32_t sy, int32_t octaves, double lacunarity, double gain,
                         double period, uint64_t seed);
/* White noise U[0,1), or quantised to `levels` levels when levels > 0. */
int mscx_gen_white(float *out, int32_t sx, int32_t sy, int32_t levels, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* MSCX_H_INCLUDED */
