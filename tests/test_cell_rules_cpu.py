"""CPU check of the gradient rules the HIP kernel runs (cell_rules.hpp,
compiled here for the host by a test-only shim) against the oracle's
raster-order assignGradients (grid_dataset.cpp:1743-1820): exhaustive on
small grids, random with ties and halos otherwise."""
import ctypes
import itertools
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.helpers import load_shim, rng_field, split_pieces, window

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM_DIR = os.path.join(HERE, "_shim")
SHIM = os.path.join(SHIM_DIR, "libcellrules.so")
DIRS = {0: (0, 0), 1: (1, 0), 2: (-1, 0), 3: (0, 1), 4: (0, -1)}


@pytest.fixture(scope="module")
def shim():
    return load_shim()


def codes_to_flags_pairs(codes, ext):
    ch, cw = codes.shape
    flags = (codes & 3).astype(np.int8)
    d = (codes >> 2) & 7
    pairs = np.full(codes.shape, 0xFFFFFFFF, np.uint32)
    ys, xs = np.nonzero(d)
    for y, x in zip(ys, xs):
        dx, dy = DIRS[int(d[y, x])]
        px, py = ext[0] + x + dx, ext[1] + y + dy
        pairs[y, x] = (px & 0xFFFF) | ((py & 0xFFFF) << 16)
    return flags, pairs


def run_case(L, win, rect, ext):
    win = np.ascontiguousarray(win, np.float32)
    cw, ch = ext[2] - ext[0] + 1, ext[3] - ext[1] + 1
    codes = np.zeros((ch, cw), np.uint8)
    r = (ctypes.c_int * 4)(*rect)
    e = (ctypes.c_int * 4)(*ext)
    assert L.cr_codes(ctypes.c_void_p(win.ctypes.data), r, e, ctypes.c_void_p(codes.ctypes.data)) == 0
    f, p = codes_to_flags_pairs(codes, ext)
    of, op, _, _ = oracle.tile_work(win, rect, ext)
    return np.array_equal(f, of) and np.array_equal(p, op)


def test_exhaustive_3x3_permutations(shim):
    rect = (0, 0, 4, 4)
    bad = 0
    for perm in itertools.permutations(range(9)):
        a = np.array(perm, np.float32).reshape(3, 3)
        if not run_case(shim, a, rect, rect):
            bad += 1
    assert bad == 0


def test_exhaustive_3x3_ties(shim):
    rect = (0, 0, 4, 4)
    for vals in itertools.product(range(3), repeat=9):
        a = np.array(vals, np.float32).reshape(3, 3)
        assert run_case(shim, a, rect, rect), vals


@pytest.mark.parametrize("kind", ["white", "ties", "const", "signedzero", "smooth"])
@pytest.mark.parametrize("tiling", [dict(levels=0), dict(levels=2), dict(lx=1, ly=1), dict(lx=2, ly=1)])
def test_random_tiles(shim, kind, tiling):
    sx, sy = 23, 19
    for seed in range(5):
        data = rng_field(kind, sx, sy, seed)
        for rect, ext in split_pieces(sx, sy, **tiling):
            assert run_case(shim, window(data, ext), rect, ext), (kind, tiling, seed, rect, ext)


def test_edge_candidate_closed_form(shim):
    """The row-step edge candidates decide an edge whose two quads are both
    usable by comparing the quads' smaller off-edge corners; every value
    pattern (with ties) of the six vertices around an edge agrees with the
    per-cell rule's quad_lt on sorted corners (grid_dataset.cpp:1354-1372,
    1519-1536)."""
    shim.cr_edge_forms_check.restype = ctypes.c_long
    assert shim.cr_edge_forms_check(6) == 0
    assert shim.cr_edge_forms_check(3) == 0
