"""Oracle -- TEST INFRASTRUCTURE ONLY.

ctypes access to liboracle.so, the CPU restatement of the reference's grid
path (oracle.cpp).  Only tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py may use it; it is the checker, never the thing measured or
shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None
P = ctypes.c_void_p


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_tile_work.restype = ctypes.c_int
        L.orc_tile_work.argtypes = [P, P, P, P, P, P, P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.orc_tile_complex.restype = ctypes.c_long
        L.orc_tile_complex.argtypes = [P, P, P, P, ctypes.c_size_t]
        L.orc_pipeline_run.restype = P
        L.orc_pipeline_run.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_pipeline_npieces.argtypes = [P]
        L.orc_pipeline_nleaves.argtypes = [P]
        L.orc_pipeline_dump.restype = ctypes.c_long
        L.orc_pipeline_dump.argtypes = [P, ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t]
        L.orc_pipeline_rect.argtypes = [P, ctypes.c_int, P]
        L.orc_pipeline_times.argtypes = [P, P]
        L.orc_pipeline_free.argtypes = [P]
        L.orc_set_plant.argtypes = [ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


class plant:
    """Context manager that runs the oracle with one planted mistake (the
    values are listed at g_plant in oracle.cpp); the pin tests use it to show
    that each check fails on a plausible error.  stop_after < 4 ends every
    pipeline after that phase (0 compute .. 4 discs), before later phases'
    own invariant checks can throw.  Not thread-safe: pipelines
    run inside it must use threads=1."""

    def __init__(self, which, stop_after=4):
        self.which = int(which)
        self.stop_after = int(stop_after)

    def __enter__(self):
        lib().orc_set_plant(self.which, self.stop_after)
        return self

    def __exit__(self, *exc):
        lib().orc_set_plant(0, 4)
        return False


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _err():
    return lib().orc_last_error().decode()


def tile_work(data, rect, ext=None):
    """GridDataset(rect, ext).init(data).work(): returns (flags, pairs, own)
    as (ch, cw) images over ext (x fastest) and the critical list."""
    ext = rect if ext is None else ext
    d = np.ascontiguousarray(data, np.float32)
    cw, ch = ext[2] - ext[0] + 1, ext[3] - ext[1] + 1
    n = cw * ch
    f = np.empty(n, np.int8)
    pr = np.empty(n, np.uint32)
    o = np.empty(n, np.uint32)
    c = np.empty(n, np.uint32)
    K = ctypes.c_size_t()
    r = (ctypes.c_int * 4)(*rect)
    e = (ctypes.c_int * 4)(*ext)
    rc = lib().orc_tile_work(_p(d), r, e, _p(f), _p(pr), _p(o), _p(c), n, ctypes.byref(K))
    if rc != 0:
        raise RuntimeError(_err())
    return f.reshape(ch, cw), pr.reshape(ch, cw), o.reshape(ch, cw), c[:K.value].copy()


def tile_complex(data, rect, ext=None):
    ext = rect if ext is None else ext
    d = np.ascontiguousarray(data, np.float32)
    r = (ctypes.c_int * 4)(*rect)
    e = (ctypes.c_int * 4)(*ext)
    n = lib().orc_tile_complex(_p(d), r, e, None, 0)
    if n < 0:
        raise RuntimeError(_err())
    buf = np.empty(n, np.uint32)
    lib().orc_tile_complex(_p(d), r, e, _p(buf), n)
    return buf


class Pipeline:
    """The GridDataManager compute order on a whole grid; dumps of every
    piece's complex after each phase (0 compute, 1 merge-up, 2 simplify,
    3 merge-down, 4 discs).  threads > 1 runs the independent pieces of each
    phase in parallel (the reference's --num-parallel); results do not
    depend on it."""

    def __init__(self, data, levels=0, lx=0, ly=0, threshold=0.0, keep=True, threads=1):
        d = np.ascontiguousarray(data, np.float32)
        sy, sx = d.shape
        self._data = d
        h = lib().orc_pipeline_run(_p(d), sx, sy, levels, lx, ly, float(threshold), 1 if keep else 0, int(threads))
        if not h:
            raise RuntimeError(_err())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_pipeline_free(self._h)
            self._h = None

    def num_pieces(self):
        return lib().orc_pipeline_npieces(self._h)

    def num_leaves(self):
        return lib().orc_pipeline_nleaves(self._h)

    def dump(self, stage, piece):
        n = lib().orc_pipeline_dump(self._h, stage, piece, None, 0)
        buf = np.empty(max(n, 0), np.uint32)
        if n > 0:
            lib().orc_pipeline_dump(self._h, stage, piece, _p(buf), n)
        return buf

    def rects(self, piece):
        r = (ctypes.c_int * 8)()
        lib().orc_pipeline_rect(self._h, piece, r)
        return tuple(r[:4]), tuple(r[4:])

    def times(self):
        t = np.zeros(5, np.float64)
        lib().orc_pipeline_times(self._h, _p(t))
        return t


def parse_dump(buf):
    """Decode a complex dump into (header, list of dicts)."""
    buf = [int(x) for x in buf]
    n = buf[0]
    hdr = {"n": n, "rect": tuple(buf[1:5]), "ext": tuple(buf[5:9])}
    cps = []
    i = 9
    for _ in range(n):
        cell, fnb, pair, flags, na, nd, nad, ndd = buf[i:i + 8]
        i += 8
        asc = buf[i:i + na]; i += na
        des = buf[i:i + nd]; i += nd
        ad = buf[i:i + nad]; i += nad
        dd = buf[i:i + ndd]; i += ndd
        cps.append(dict(cell=cell, fn=fnb, pair=pair, flags=flags, asc=asc, des=des, asc_disc=ad, des_disc=dd))
    return hdr, cps
