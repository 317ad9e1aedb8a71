"""Shared helpers for the tests (tiling geometry, inputs, comparisons)."""
import numpy as np


def split_pieces(sx, sy, levels=0, lx=0, ly=0):
    """Leaf (rect, ext) list in piece order, following the reference's
    createPieces_quadtree split/halo rule (grid_datamanager.cpp:53-89); with
    lx/ly the split axis alternates as in the oracle's 2-D tiling."""
    two_d = lx > 0 or ly > 0
    total = lx + ly if two_d else levels
    out = []

    def rec(r, e, level, lx, ly):
        if level == 0:
            out.append((tuple(r), tuple(e)))
            return
        dim = 1
        if two_d:
            dim = 1 if (ly >= lx and ly > 0) else 0
        s = [r[2] - r[0], r[3] - r[1]]
        tr1 = [r[2], r[3]]
        bl1 = [r[0], r[1]]
        tr1[dim] -= 2 * (s[dim] // 4)
        bl1[dim] = tr1[dim]
        r1 = (r[0], r[1], tr1[0], tr1[1])
        r2 = (bl1[0], bl1[1], r[2], r[3])
        tr2 = [e[2], e[3]]
        bl2 = [e[0], e[1]]
        tr2[dim] = tr1[dim] + 2
        bl2[dim] = bl1[dim] - 2
        e1 = (e[0], e[1], tr2[0], tr2[1])
        e2 = (bl2[0], bl2[1], e[2], e[3])
        nlx, nly = lx, ly
        if two_d:
            if dim == 1:
                nly -= 1
            else:
                nlx -= 1
        rec(r1, e1, level - 1, nlx, nly)
        rec(r2, e2, level - 1, nlx, nly)

    r = (0, 0, 2 * (sx - 1), 2 * (sy - 1))
    rec(r, r, total, lx, ly)
    return out


def window(data, ext):
    """The ext window (vertex values) of a full grid."""
    x0, y0, x1, y1 = ext[0] // 2, ext[1] // 2, ext[2] // 2, ext[3] // 2
    return np.ascontiguousarray(data[y0:y1 + 1, x0:x1 + 1])


def rng_field(kind, sx, sy, seed):
    r = np.random.default_rng(seed)
    if kind == "white":
        return r.random((sy, sx), dtype=np.float32)
    if kind == "ties":
        return r.integers(0, 3, (sy, sx)).astype(np.float32)
    if kind == "const":
        return np.zeros((sy, sx), np.float32)
    if kind == "signedzero":
        a = r.integers(-1, 2, (sy, sx)).astype(np.float32)
        a[a == 0] = np.where(r.random(int((a == 0).sum())) < 0.5, -0.0, 0.0)
        return a
    if kind == "smooth":
        y, x = np.mgrid[0:sy, 0:sx]
        return (np.sin(x / 3.0) * np.cos(y / 4.0) + 0.01 * r.random((sy, sx))).astype(np.float32)
    if kind == "denormal":
        return (r.random((sy, sx)) * 1e-39).astype(np.float32)
    raise ValueError(kind)


def load_shim():
    """Builds (once per process) and loads the test-only CPU build of the
    gradient rules (tests/cell_rules_shim.cpp)."""
    import ctypes
    import os
    import subprocess
    global _SHIM
    if _SHIM is None:
        here = os.path.dirname(os.path.abspath(__file__))
        d = os.path.join(here, "_shim")
        os.makedirs(d, exist_ok=True)
        so = os.path.join(d, "libcellrules.so")
        subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so,
                        os.path.join(here, "cell_rules_shim.cpp")], check=True)
        L = ctypes.CDLL(so)
        L.cr_codes.argtypes = [ctypes.c_void_p] * 4
        L.cr_codes_region.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 4 + [ctypes.c_void_p]
        _SHIM = L
    return _SHIM


_SHIM = None
