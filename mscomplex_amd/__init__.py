"""mscomplex_amd -- MI355X Morse-Smale complex engine for 2D float32 grids.

Python mirror of the reference's grid path (dmsc): GridDataset,
GridMSComplex and the GridDataManager compute order, backed by libmscx.so
(hand-written HIP kernels for gfx950 plus a C++ host runtime).  There is no
CPU fallback: every compute call goes through the HIP library.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import MscxError, check, lib, ptr

__all__ = ["Batch", "DeviceBuffer", "profile", "profile_read", "GridDataset", "GridMSComplex", "GridRun", "work_tiles", "gen_gaussians", "gen_value_noise",
           "gen_white", "MscxError", "unpack", "pack", "grid_rect"]


def pack(x, y):
    return (int(x) & 0xFFFF) | ((int(y) & 0xFFFF) << 16)


def unpack(p):
    p = int(p)
    x, y = p & 0xFFFF, p >> 16
    return (x - 0x10000 if x >= 0x8000 else x, y - 0x10000 if y >= 0x8000 else y)


def grid_rect(size_x, size_y):
    """Rect of a whole size_x x size_y vertex grid in cell coordinates."""
    return (0, 0, 2 * (size_x - 1), 2 * (size_y - 1))


# ------------------------------------------------------------------ inputs

def gen_gaussians(sx, sy, n=32, noise=1e-3, seed=0):
    out = np.empty((sy, sx), np.float32)
    check(lib().mscx_gen_gaussians(ptr(out), sx, sy, n, noise, seed))
    return out


def gen_value_noise(sx, sy, octaves=8, lacunarity=2.0, gain=0.5, period=None, seed=0):
    if period is None:
        period = max(sx, sy) / 4.0
    out = np.empty((sy, sx), np.float32)
    check(lib().mscx_gen_value_noise(ptr(out), sx, sy, octaves, lacunarity, gain, float(period), seed))
    return out


def gen_white(sx, sy, levels=0, seed=0):
    out = np.empty((sy, sx), np.float32)
    check(lib().mscx_gen_white(ptr(out), sx, sy, levels, seed))
    return out


# ------------------------------------------------------------------ device buffers

class DeviceBuffer:
    """A HBM allocation owned by libmscx (keeps grids resident on the GPU)."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().mscx_dev_alloc(nbytes, ctypes.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().mscx_dev_upload(ctypes.c_void_p(b.ptr), ptr(a), a.nbytes))
        return b

    def upload(self, a):
        """Overwrite the start of the buffer with the bytes of `a`."""
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        check(lib().mscx_dev_upload(ctypes.c_void_p(self.ptr), ptr(a), a.nbytes))

    def download(self, out):
        check(lib().mscx_dev_download(ptr(out), ctypes.c_void_p(self.ptr), out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().mscx_dev_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def profile(enable=True):
    """Enable (and reset) per-stage HIP-event timing inside the library."""
    check(lib().mscx_profile(1 if enable else 0))


def set_graphs(on=True):
    """Submit the tile stages that need no host decision as a replayed HIP
    graph (True, the default) or as direct launches; returns the old setting."""
    prev = ctypes.c_int(0)
    check(lib().mscx_set_graphs(1 if on else 0, ctypes.byref(prev)))
    return bool(prev.value)


def graph_stats():
    """(builds, replays) of the stage graph since initialisation."""
    c, r = ctypes.c_uint64(0), ctypes.c_uint64(0)
    check(lib().mscx_graph_stats(ctypes.byref(c), ctypes.byref(r)))
    return int(c.value), int(r.value)


def set_block_rows(table=None):
    """Block tile height (32 or 16 vertex rows) of the gradient, trace and
    owner kernels by tile height: table = [(max_rows, rows), ...] with
    increasing bounds; a tile of vy vertex rows takes the first entry with
    vy <= max_rows, 32 rows when none matches.  None restores the default.
    Applies to tiles created afterwards; results do not depend on it."""
    table = list(table or [])
    b = np.array([t[0] for t in table], np.int32)
    r = np.array([t[1] for t in table], np.int32)
    check(lib().mscx_set_block_rows(len(table), ptr(b) if table else None, ptr(r) if table else None))


def disc_trace_stats():
    """(disc slots traced on the GPU, slots traced on the host, calls whose
    GPU lists were dropped, GPU slots traced by a whole wavefront, longest
    list traced on the GPU) since the library loaded."""
    o = np.zeros(5, np.uint64)
    check(lib().mscx_disc_trace_stats(ptr(o)))
    return tuple(int(v) for v in o)


def merge_stats():
    """(merges run on the GPU, grids whose merge tree fell back to the host
    path, host passes over shared lines) since the library loaded."""
    o = np.zeros(3, np.uint64)
    check(lib().mscx_merge_stats(ptr(o)))
    return tuple(int(v) for v in o)


STAGES = ("gradient", "scan", "compact", "trace", "owner", "incidence", "host_complex",
          "disc_prep", "disc_tables", "disc_label_wait_thread_time", "disc_trace", "disc_lists_thread_time",
          "disc_group_thread_time", "disc_group", "disc_codes", "leaf_copies")


def profile_read():
    ms = np.zeros(len(STAGES), np.float64)
    cnt = np.zeros(len(STAGES), np.uint32)
    check(lib().mscx_profile_read(ptr(ms), ptr(cnt), len(STAGES)))
    return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(STAGES)}


# ------------------------------------------------------------------ tiles

class GridDataset:
    """One tile (rect + ext window) on the GPU; mirrors dmsc GridDataset
    (grid_dataset.h:35-240).  rect/ext are (x0, y0, x1, y1) cell coords."""

    def __init__(self, rect, ext=None):
        ext = rect if ext is None else ext
        self.rect = tuple(int(v) for v in rect)
        self.ext = tuple(int(v) for v in ext)
        h = ctypes.c_void_p()
        r = (ctypes.c_int32 * 4)(*self.rect)
        e = (ctypes.c_int32 * 4)(*self.ext)
        check(lib().mscx_tile_create(r, e, ctypes.byref(h)))
        self._h = h
        self.cw = self.ext[2] - self.ext[0] + 1
        self.ch = self.ext[3] - self.ext[1] + 1
        self.vx = self.cw // 2 + 1
        self.vy = self.ch // 2 + 1

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            try:
                lib().mscx_tile_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def init(self, data):
        """GridDataset::init: the ext window of vertex values (vy x vx, x fastest)."""
        a = np.ascontiguousarray(data, np.float32).reshape(-1)
        check(lib().mscx_tile_upload(self._h, ptr(a), a.size))

    def bind_device(self, dev_ptr, row_pitch):
        check(lib().mscx_tile_bind_device(self._h, ctypes.c_void_p(dev_ptr), row_pitch))

    def work(self):
        check(lib().mscx_tile_work(self._h))

    def num_critical(self):
        k = ctypes.c_uint32()
        check(lib().mscx_tile_num_critical(self._h, ctypes.byref(k)))
        return k.value

    def arrays(self, flags=True, pairs=True, own=True, crit=True):
        """(flags int8, pairs u32, own u32) as (ch, cw) images over ext, and
        the critical list (K packed ids)."""
        n = self.cw * self.ch
        f = np.empty(n, np.int8) if flags else None
        p = np.empty(n, np.uint32) if pairs else None
        o = np.empty(n, np.uint32) if own else None
        c = np.empty(self.num_critical(), np.uint32) if crit else None
        check(lib().mscx_tile_download(self._h, ptr(f), ptr(p), ptr(o), ptr(c)))
        shp = (self.ch, self.cw)
        return (None if f is None else f.reshape(shp), None if p is None else p.reshape(shp),
                None if o is None else o.reshape(shp), c)

    def codes(self):
        c = np.empty(self.cw * self.ch, np.uint8)
        check(lib().mscx_tile_download_codes(self._h, ptr(c)))
        return c.reshape(self.ch, self.cw)

    def critical_data(self):
        k = self.num_critical()
        fn = np.empty(k, np.float32)
        conn = np.empty(4 * k, np.uint32)
        pr = np.empty(k, np.uint32)
        check(lib().mscx_tile_download_critical(self._h, ptr(fn), ptr(conn), ptr(pr)))
        return fn, conn.reshape(k, 4), pr

    def fill_discs(self, msc):
        """GridDataset::postMergeFillDiscs into `msc`."""
        check(lib().mscx_tile_fill_discs(self._h, msc.handle))


def work_tiles(tiles):
    """Batched work() over a list of GridDataset (one launch per stage)."""
    arr = (ctypes.c_void_p * len(tiles))(*[t.handle.value for t in tiles])
    check(lib().mscx_tiles_work(arr, len(tiles)))


# ------------------------------------------------------------------ complexes

class GridMSComplex:
    """Host-resident MS complex; mirrors dmsc GridMSComplex."""

    def __init__(self, handle, owned=True, owner=None):
        self._h = handle
        self._owned = owned
        # a borrowed complex keeps the object that owns it alive
        self._owner = owner

    def __del__(self):
        if not getattr(self, "_owned", False):
            return
        h = getattr(self, "_h", None)
        if h and h.value:
            try:
                lib().mscx_complex_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    @classmethod
    def from_dataset(cls, ds):
        """GridDataset::writeout_connectivity into a new complex."""
        h = ctypes.c_void_p()
        check(lib().mscx_complex_from_tile(ds.handle, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_critical(cls, rect, ext, crit, fn, conn4, pair_idx):
        """writeout_connectivity_ocl counterpart: a complex from per-critical
        data (raster-ordered packed ids, values, 4 incident owner indices per
        cell, pair indices)."""
        crit = np.ascontiguousarray(crit, np.uint32)
        fn = np.ascontiguousarray(fn, np.float32)
        conn4 = np.ascontiguousarray(conn4, np.uint32).reshape(-1)
        pair_idx = np.ascontiguousarray(pair_idx, np.uint32)
        h = ctypes.c_void_p()
        r = (ctypes.c_int32 * 4)(*rect)
        e = (ctypes.c_int32 * 4)(*ext)
        check(lib().mscx_complex_from_critical(r, e, crit.size, ptr(crit), ptr(fn), ptr(conn4), ptr(pair_idx),
                                               ctypes.byref(h)))
        return cls(h)

    def fill_discs_host(self, rect, ext, codes, own):
        """postMergeFillDiscs from host copies of a tile's codes and owners."""
        codes = np.ascontiguousarray(codes, np.uint8)
        own = np.ascontiguousarray(own, np.uint32).copy()
        r = (ctypes.c_int32 * 4)(*rect)
        e = (ctypes.c_int32 * 4)(*ext)
        check(lib().mscx_complex_fill_discs_host(self._h, r, e, ptr(codes), ptr(own)))
        return own

    @classmethod
    def merge_up(cls, a, b):
        h = ctypes.c_void_p()
        check(lib().mscx_complex_merge_up(a.handle, b.handle, ctypes.byref(h)))
        return cls(h)

    def merge_down(self, a, b):
        check(lib().mscx_complex_merge_down(self._h, a.handle, b.handle))

    def simplify_un_simplify(self, threshold):
        check(lib().mscx_complex_simplify_un_simplify(self._h, float(threshold)))

    def write_discs(self, prefix):
        check(lib().mscx_complex_write_discs(self._h, prefix.encode()))

    def save(self, path):
        """Host archive of the complex (write_msgraph_to_archive counterpart)."""
        check(lib().mscx_complex_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path):
        """A complex from an archive written by save() (read_msgraph_from_archive)."""
        h = ctypes.c_void_p()
        check(lib().mscx_complex_load(os.fsencode(path), ctypes.byref(h)))
        return cls(h)

    def num_cps(self):
        n = ctypes.c_uint32()
        check(lib().mscx_complex_num_cps(self._h, ctypes.byref(n)))
        return n.value

    def dump(self):
        """Canonical uint32 dump (format in include/mscx.h)."""
        n = ctypes.c_size_t()
        check(lib().mscx_complex_dump(self._h, None, 0, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint32)
        check(lib().mscx_complex_dump(self._h, ptr(buf), buf.size, ctypes.byref(n)))
        return buf


class _ViewComplex(GridMSComplex):
    """A piece of a served request's grid: valid while the request is."""

    @property
    def _h(self):
        self._owner._handle()  # raises once the request is released
        return self.__dict__["_hh"]

    @_h.setter
    def _h(self, v):
        self.__dict__["_hh"] = v


class _GridView:
    """Borrowed view of one grid of a batch or of a served request (same
    accessors as GridRun).  It keeps its owner alive; a served request's
    view becomes invalid when the request is released."""

    def __init__(self, handle, owner):
        self._h = handle
        self._owner = owner

    def _handle(self):
        if self._h is None:
            raise MscxError(-5, "the request of this view has been released")
        return self._h

    def num_pieces(self):
        n = ctypes.c_int32()
        check(lib().mscx_grid_num_pieces(self._handle(), ctypes.byref(n)))
        return n.value

    def piece(self, i):
        h = ctypes.c_void_p()
        check(lib().mscx_grid_piece(self._handle(), i, ctypes.byref(h)))
        return _ViewComplex(h, owned=False, owner=self)

    def times(self):
        t = np.zeros(5, np.float64)
        check(lib().mscx_grid_times(self._handle(), ptr(t)))
        return t

    def checksum(self):
        """FNV-1a of every piece's dump words (mscx_grid_checksum)."""
        h = ctypes.c_uint64()
        check(lib().mscx_grid_checksum(self._handle(), ctypes.byref(h)))
        return h.value


class Batch:
    """Many grids through the pipeline together (mscx_batch_run).  grids:
    list of 2-D float32 arrays (or, with device=True, (ptr, sx, sy) tuples)."""

    def __init__(self, grids, lx, ly, thresholds, discs=True, device=False):
        n = len(grids)
        lx = np.broadcast_to(np.asarray(lx, np.int32), (n,)).copy()
        ly = np.broadcast_to(np.asarray(ly, np.int32), (n,)).copy()
        th = np.broadcast_to(np.asarray(thresholds, np.float64), (n,)).copy()
        sx = np.empty(n, np.int32)
        sy = np.empty(n, np.int32)
        ptrs = (ctypes.c_void_p * max(n, 1))()
        keep = []
        for i, g in enumerate(grids):
            if device:
                ptrs[i], sx[i], sy[i] = g[0], g[1], g[2]
            else:
                a = np.ascontiguousarray(g, np.float32)
                keep.append(a)
                sy[i], sx[i] = a.shape
                ptrs[i] = a.ctypes.data
        flags = (0 if discs else 2) | (1 if device else 0)
        h = ctypes.c_void_p()
        check(lib().mscx_batch_run(n, ptrs, ptr(sx), ptr(sy), ptr(lx), ptr(ly), ptr(th), flags, ctypes.byref(h)))
        self._h = h
        self.n = n

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            try:
                lib().mscx_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def status(self, i):
        g = ctypes.c_void_p()
        st = ctypes.c_int32()
        check(lib().mscx_batch_grid(self._h, i, ctypes.byref(g), ctypes.byref(st)))
        return st.value, lib().mscx_batch_error(self._h, i).decode()

    def grid(self, i):
        g = ctypes.c_void_p()
        st = ctypes.c_int32()
        check(lib().mscx_batch_grid(self._h, i, ctypes.byref(g), ctypes.byref(st)))
        if st.value != 0:
            raise MscxError(st.value, lib().mscx_batch_error(self._h, i).decode())
        return _GridView(g, self)


class Server:
    """Continuous-batching request loop (mscx_server_*): requests are
    admitted against an HBM budget, the GPU stages of every grid admitted in
    one round run together, host phases overlap the next round.  A nonzero
    key keeps the grid resident for repeat requests at other thresholds."""

    def __init__(self, hbm_budget, host_threads=0):
        h = ctypes.c_void_p()
        check(lib().mscx_server_create(int(hbm_budget), int(host_threads), ctypes.byref(h)))
        self._h = h
        self._views = {}  # ticket -> views handed out by wait(), invalidated by release()

    def close(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            check(lib().mscx_server_destroy(h))
            self._h = None
            for vs in self._views.values():
                for v in vs:
                    v._h = None
            self._views.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_work_limit(self, factor):
        """Simplification work limit of later requests (x their initial heap
        entries; 0 = none, the reference's unbounded behaviour)."""
        check(lib().mscx_server_set_work_limit(self._h, float(factor)))

    def submit(self, data, lx=0, ly=0, threshold=0.0, key=0, discs=True, size=None):
        """data: 2-D float32 array, or None for a resident key (then size=(sx, sy))."""
        t = ctypes.c_uint64()
        if data is None:
            sx, sy = size
            p = None
        else:
            data = np.ascontiguousarray(data, np.float32)
            sy, sx = data.shape
            p = ptr(data)
        check(lib().mscx_server_submit(self._h, p, sx, sy, lx, ly, float(threshold), int(key),
                                       0 if discs else 2, ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """(status, error, latency seconds, grid view or None)."""
        g = ctypes.c_void_p()
        st = ctypes.c_int32()
        lat = ctypes.c_double()
        check(lib().mscx_server_wait(self._h, ticket, float(timeout), ctypes.byref(g), ctypes.byref(st),
                                     ctypes.byref(lat)))
        err = lib().mscx_server_error(self._h, ticket).decode()
        view = None
        if st.value == 0 and g.value:
            view = _GridView(g, self)
            self._views.setdefault(ticket, []).append(view)
        return st.value, err, lat.value, view

    def release(self, ticket):
        check(lib().mscx_server_release(self._h, ticket))
        for v in self._views.pop(ticket, []):
            v._h = None

    def stats(self):
        o = np.zeros(8, np.uint64)
        check(lib().mscx_server_stats(self._h, ptr(o)))
        keys = ("done", "rounds", "repeats", "evicted", "resident", "device_bytes", "budget", "queued")
        return {k: int(v) for k, v in zip(keys, o)}


RUN_DEVICE_MERGE, RUN_HOST_MERGE = 8, 16


def _merge_flags(merge):
    if merge is None:
        return 0
    if merge == "device":
        return RUN_DEVICE_MERGE
    if merge == "host":
        return RUN_HOST_MERGE
    raise ValueError("merge must be None, 'device' or 'host'")


class GridRun:
    """A whole GridDataManager compute run over one grid (mscx_grid_run)."""

    def __init__(self, data, levels=1, lx=0, ly=0, threshold=0.0, discs=True, device_ptr=None, size=None,
                 out_of_core=False, merge=None):
        """merge: None lets the library choose where the merge tree runs,
        "device" keeps the pieces' complexes in HBM and merges them with
        the GPU kernels, "host" merges host complexes (the results are the
        same)."""
        h = ctypes.c_void_p()
        flags = (0 if discs else 2) | (4 if out_of_core else 0) | _merge_flags(merge)
        if device_ptr is not None:
            sx, sy = size
            flags |= 1
            p = ctypes.c_void_p(device_ptr)
        else:
            data = np.ascontiguousarray(data, np.float32)
            sy, sx = data.shape
            p = ptr(data)
        check(lib().mscx_grid_run(p, sx, sy, levels, lx, ly, float(threshold), flags, ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_file(cls, path, size, levels=1, lx=0, ly=0, threshold=0.0, num_parallel=0, discs=True,
                  disc_prefix=None, out_of_core=False, merge=None):
        """The reference driver over a raw float32 grid file (x fastest) of
        size (sx, sy) (mscx_grid_run_file): GridDataManager(filename, sx, sy,
        levels, ..., simp_tresh, ..., num_parallel).  disc_prefix, when
        given, receives every leaf's disc files (the reference driver uses
        "dp__disc_" in its working directory).  out_of_core keeps idle pieces
        serialised (the reference's --out-of-core)."""
        sx, sy = size
        h = ctypes.c_void_p()
        pre = None if disc_prefix is None else os.fsencode(disc_prefix)
        check(lib().mscx_grid_run_file(os.fsencode(path), sx, sy, levels, lx, ly, float(threshold), int(num_parallel),
                                       (0 if discs else 2) | (4 if out_of_core else 0) | _merge_flags(merge),
                                       pre, ctypes.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            try:
                lib().mscx_grid_destroy(h)
            except Exception:
                pass
            self._h = None

    def num_pieces(self):
        n = ctypes.c_int32()
        check(lib().mscx_grid_num_pieces(self._h, ctypes.byref(n)))
        return n.value

    def piece(self, i):
        h = ctypes.c_void_p()
        check(lib().mscx_grid_piece(self._h, i, ctypes.byref(h)))
        return GridMSComplex(h, owned=False, owner=self)

    def checksum(self):
        """FNV-1a of every piece's dump words (mscx_grid_checksum)."""
        h = ctypes.c_uint64()
        check(lib().mscx_grid_checksum(self._h, ctypes.byref(h)))
        return h.value

    def piece_num_cps(self, i):
        """Critical points of piece i (no host copy of a device-resident piece)."""
        n = ctypes.c_uint32()
        check(lib().mscx_grid_piece_num_cps(self._h, i, ctypes.byref(n)))
        return n.value

    def times(self):
        t = np.zeros(5, np.float64)
        check(lib().mscx_grid_times(self._h, ptr(t)))
        return t


# ------------------------------------------------------------- workloads

def uneven_mix(n, side_lo, side_hi, seed, aspect_max=4.0, thresholds=None):
    """n grids with sides log-uniform in [side_lo, side_hi] (the long side),
    aspect ratio uniform in [1, aspect_max] with random orientation, half
    Gaussian sums (32 Gaussians, +/-1e-3 noise) and half 8-octave value
    noise; returns a list of (array, threshold)."""
    r = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(round(np.exp(r.uniform(np.log(side_lo), np.log(side_hi)))))
        short = max(2, int(round(side / r.uniform(1.0, aspect_max))))
        sx, sy = (side, short) if r.random() < 0.5 else (short, side)
        s = int(r.integers(1 << 62))
        g = gen_gaussians(sx, sy, 32, 1e-3, seed=s) if i % 2 else gen_value_noise(sx, sy, seed=s, period=max(sx, sy) / 4)
        th = float(thresholds(r)) if callable(thresholds) else float(thresholds or 0.0)
        out.append((g, th))
    return out


def _pick_tiling(res, margin):
    """From [(seconds, tiling)] in candidate order (coarsest first): the
    first, replaced by a later one only when faster by more than margin."""
    choice = None
    for el, tl in res:
        if choice is None or el < choice[0] * (1.0 - margin):
            choice = (el, tl)
    return choice[1] if choice else (0, 0)


def autotune_block_rows(grids, thresholds, tiling_policy=None, classes=(128, 512, 1 << 30), candidates=(32, 16),
                        reps=2, margin=0.03):
    """Kernel block tile height per tile-height class for batched runs,
    chosen by timing the GPU stages.

    Grids are classed by the vertex rows of their leaf tiles under
    tiling_policy ((sx, sy) -> (lx, ly); default no tiling).  For each class
    and each candidate height, the class's grids run as one batch (without
    manifolds) with per-stage HIP-event timing; the GPU stage time (gradient,
    scan, compaction, trace, owners, incidence) is the best of `reps` runs.
    The first candidate (the library's default) is kept unless another is
    faster by more than `margin`.  The chosen table is installed with
    set_block_rows; returns (table, measured ms per (class bound, height))."""
    tiling_policy = tiling_policy or (lambda sx, sy: (0, 0))
    by_class = {}
    for g, t in zip(grids, thresholds):
        sy, sx = g.shape
        lx, ly = tiling_policy(sx, sy)
        rows = (sy - 1) // (1 << ly) + 3
        bound = next(c for c in classes if rows < c)
        by_class.setdefault(bound, []).append((g, t, lx, ly))
    meas, table = {}, []
    try:
        for bound in classes:
            members = by_class.get(bound)
            if not members:
                continue
            res = []
            for rows in candidates:
                set_block_rows([(1 << 30, rows)])
                best = float("inf")
                for _ in range(reps):
                    profile(True)
                    Batch([m[0] for m in members], [m[2] for m in members], [m[3] for m in members],
                          [m[1] for m in members], discs=False)
                    st = profile_read()
                    best = min(best, sum(st[k][0] for k in STAGES[:6]))
                meas[(bound, rows)] = best
                res.append((best, rows))
            choice = res[0]
            for el, rows in res[1:]:
                if el < choice[0] * (1.0 - margin):
                    choice = (el, rows)
            table.append((bound - 1, choice[1]))
    finally:
        profile(False)
    # drop trailing classes that keep the default height
    while table and table[-1][1] == candidates[0]:
        table.pop()
    set_block_rows(table)
    return table, meas


def autotune_tiling(grids, thresholds, candidates=((0, 0), (1, 1), (2, 2)), classes=(256, 1024, 1 << 30),
                    reps=1, margin=0.05):
    """Leaf tiling (lx, ly) per size class for batched runs, chosen by timing.

    For each size class (grids whose larger side is below the class bound),
    every candidate tiling runs the class's grids as one batch.  Candidates
    are taken in order (coarsest first) and a later one replaces the current
    choice only when it is faster by more than `margin` (a fraction): timing
    noise on a small sample does not pick a finer tiling, whose merge tree
    costs host time in the full batch.  Returns (policy, table):
    policy(sx, sy) -> (lx, ly), and the measured seconds per (class bound,
    candidate)."""
    import time as _time
    table = {}
    best = {}
    for bound in classes:
        members = [(g, t) for g, t in zip(grids, thresholds)
                   if max(g.shape) < bound and all(max(g.shape) >= b for b in classes if b < bound)]
        if not members:
            continue
        res = []
        for lx, ly in candidates:
            if any(min(g.shape) < 8 * (1 << max(lx, ly)) for g, _ in members):
                continue
            el = float("inf")
            for _ in range(reps):
                t0 = _time.perf_counter()
                Batch([g for g, _ in members], lx, ly, [t for _, t in members])
                el = min(el, _time.perf_counter() - t0)
            table[(bound, (lx, ly))] = el
            res.append((el, (lx, ly)))
        best[bound] = _pick_tiling(res, margin)

    def policy(sx, sy):
        for bound in classes:
            if max(sx, sy) < bound:
                return best.get(bound, (0, 0))
        return (0, 0)

    return policy, table
