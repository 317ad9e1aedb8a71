"""Math-fixed checks of Morse-Smale complex dumps (tests/pins.cpp).

Test infrastructure: the same checks run on the oracle's complexes (CPU
tests) and on the HIP pipeline's (GPU tests, also at sizes the oracle cannot
reach).  None of them restates the reference's algorithm; each is a fact of
the mathematics that a wrong arc rewiring, dedup or rebuild breaks.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None
P = ctypes.c_void_p

F_CANCELLED, F_STRANG, F_BCANC = 1, 2, 4


def lib():
    global _L
    if _L is None:
        d = os.path.join(HERE, "_shim")
        os.makedirs(d, exist_ok=True)
        so = os.path.join(d, "libpins.so")
        src = os.path.join(HERE, "pins.cpp")
        if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
            tmp = so + f".{os.getpid()}.tmp"
            subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", tmp, src], check=True)
            os.replace(tmp, so)
        L = ctypes.CDLL(so)
        L.pin_nodes.restype = ctypes.c_long
        L.pin_nodes.argtypes = [P, ctypes.c_size_t, P, P, P, P, P]
        L.pin_canon.restype = ctypes.c_long
        L.pin_canon.argtypes = [P, ctypes.c_size_t, P, P, ctypes.c_size_t]
        L.pin_homology.argtypes = [P, ctypes.c_size_t, P, P]
        L.pin_discs.argtypes = [P, ctypes.c_size_t, P, P]
        _L = L
    return _L


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _u32(d):
    return np.ascontiguousarray(d, np.uint32)


def nodes(dump):
    """Per-node arrays of a dump: dict(cell, fn, pair, flags, counts (K, 4)),
    plus rect and ext."""
    d = _u32(dump)
    k = lib().pin_nodes(_p(d), d.size, None, None, None, None, None)
    assert k >= 0, "malformed dump"
    a = {x: np.empty(k, np.uint32) for x in ("cell", "fn", "pair", "flags")}
    cnt = np.empty((k, 4), np.uint32)
    lib().pin_nodes(_p(d), d.size, _p(a["cell"]), _p(a["fn"]), _p(a["pair"]), _p(a["flags"]), _p(cnt))
    a["counts"] = cnt
    a["rect"] = tuple(int(np.int32(v)) for v in d[1:5].view(np.int32))
    a["ext"] = tuple(int(v) for v in d[5:9].view(np.int32))
    return a


def dims(cells):
    c = np.asarray(cells, np.uint32)
    return (c & 1) + ((c >> 16) & 1)


def canon(dump, mask=None):
    """Canonical words of the selected nodes (default: not cancelled), by
    cell id, arcs as sorted cell-id multisets.  Raises if an arc of a
    selected node leads outside the selection."""
    d = _u32(dump)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    n = lib().pin_canon(_p(d), d.size, _p(m), None, 0)
    assert n != -1, "malformed dump"
    assert n != -2, "an arc of a selected node leads to an unselected node"
    out = np.empty(n, np.uint32)
    lib().pin_canon(_p(d), d.size, _p(m), _p(out), n)
    return out


HOMOLOGY_KEYS = ("min", "saddle", "max", "rank_d1", "rank_d2", "odd_d1d2", "arcs_out", "bad_des_weight",
                 "bad_asc_weight", "bad_dim", "unmirrored")


def homology(dump, mask):
    d = _u32(dump)
    m = np.ascontiguousarray(mask, np.uint8)
    out = np.zeros(11, np.int64)
    assert lib().pin_homology(_p(d), d.size, _p(m), _p(out)) == 0, "malformed dump"
    return dict(zip(HOMOLOGY_KEYS, (int(v) for v in out)))


def assert_disc_homology(dump, mask, what=""):
    """The selected nodes form the Morse complex of a disc over Z/2: d1 d2 = 0,
    rank d1 = #min - 1, rank d2 = #max, every saddle has two descending
    paths and two ascending ones (one on the boundary), arcs are mirrored
    and join adjacent dimensions."""
    h = homology(dump, mask)
    bad = {k: h[k] for k in HOMOLOGY_KEYS[5:] if h[k]}
    assert not bad, (what, h)
    assert h["rank_d1"] == h["min"] - 1, (what, h)
    assert h["rank_d2"] == h["max"], (what, h)
    assert h["min"] - h["saddle"] + h["max"] == 1, (what, h)
    return h


DISC_KEYS = ("vert_missing", "vert_twice", "vert_foreign", "quad_missing", "quad_twice", "quad_foreign",
             "dead_disc", "unsorted")


def discs(dump, rect):
    d = _u32(dump)
    r = np.asarray(rect, np.int32)
    out = np.zeros(8, np.int64)
    assert lib().pin_discs(_p(d), d.size, _p(r), _p(out)) == 0, "malformed dump"
    return dict(zip(DISC_KEYS, (int(v) for v in out)))


def assert_disc_partition(dump, rect, what=""):
    r = discs(dump, rect)
    assert not any(r.values()), (what, r)


def survivors(dump):
    """Nodes that survive postMergeFillDiscs' test: not boundary-cancellable."""
    n = nodes(dump)
    return (n["flags"] & F_BCANC) == 0


def cells_in(cells, rect):
    c = np.asarray(cells, np.uint32)
    x = (c & 0xFFFF).astype(np.int32)
    y = (c >> 16).astype(np.int32)
    return (x >= rect[0]) & (x <= rect[2]) & (y >= rect[1]) & (y <= rect[3])


def without_fn(dump):
    """The dump with every function-value word zeroed (compares complexes of
    fields that differ by a shift)."""
    d = _u32(dump).copy()
    n = nodes(d)
    rec = 8 + n["counts"].astype(np.int64).sum(axis=1)
    start = 9 + np.concatenate(([0], np.cumsum(rec)[:-1]))
    d[start + 1] = 0
    return d
