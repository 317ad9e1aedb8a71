"""Worst-case configuration at full size: 16384x16384 white noise U[0,1)
(dense critical cells) and its 256-level quantisation (ties on plateaus),
through the GPU tile stages (gradient, collation, trace, incidence) on one
whole-grid tile and on the 8x8 tiling.

The oracle does not finish the whole grid at this size, but it does finish
single leaves of the 8x8 tiling: three leaves of each field (two corners and
an interior leaf, 2049 x 2049 vertices with halos) are compared with the
oracle's GridDataset work word for word (flags, pairs, owners, critical
list).  Beyond those, parity rests on properties that hold at any size and
pin each stage against an independent computation:
  * gradient: every pairing is a mutual involution, and sampled windows
    (random interiors, the corners and the tile sides) equal the per-cell
    rules evaluated on the CPU (checked exhaustively against the oracle in
    test_cell_rules_cpu.py);
  * collation: K equals the number of critical flags and the critical list
    is their raster order; cell values of sampled critical cells equal the
    closed form of get_cell_fn (ordered float sum, then the division);
  * trace: owners of sampled cells equal a sequential chase of the codes;
  * incidence: sampled saddles' entries equal the owners of their facets and
    cofacets looked up in the critical list;
  * the whole-grid tile has Euler characteristic #min - #saddle + #max = 1.
"""
import ctypes

import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from tests.helpers import load_shim, split_pieces, window

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

SIDE = 16384
NONE = 0xFFFFFFFF
DX = np.array([0, 1, -1, 0, 0])
DY = np.array([0, 0, 0, 1, -1])


def _pack(x, y):
    return (np.asarray(x, np.uint32) & 0xFFFF) | ((np.asarray(y, np.uint32) & 0xFFFF) << 16)


def _involution(codes):
    d = (codes >> 2) & 7
    assert np.all(d[:, 1:][d[:, :-1] == 1] == 2), "+x pairing not mutual"
    assert np.all(d[:, :-1][d[:, 1:] == 2] == 1), "-x pairing not mutual"
    assert np.all(d[1:, :][d[:-1, :] == 3] == 4), "+y pairing not mutual"
    assert np.all(d[:-1, :][d[1:, :] == 4] == 3), "-y pairing not mutual"
    assert not np.any(d[:, -1] == 1) and not np.any(d[:, 0] == 2), "pair outside ext"
    assert not np.any(d[-1, :] == 3) and not np.any(d[0, :] == 4), "pair outside ext"


def _windows(codes, win, rect, ext, rng, label, n=6, w=96):
    shim = load_shim()
    ch, cw = codes.shape
    spots = [(0, 0), (cw - w, 0), (0, ch - w), (cw - w, ch - w), (cw // 2, 0), (0, ch // 2)]
    spots += [(int(rng.integers(0, cw - w)), int(rng.integers(0, ch - w))) for _ in range(n)]
    win = np.ascontiguousarray(win, np.float32)
    r = (ctypes.c_int * 4)(*rect)
    e = (ctypes.c_int * 4)(*ext)
    for x0, y0 in spots:
        out = np.empty((w, w), np.uint8)
        assert shim.cr_codes_region(win.ctypes.data, r, e, x0, y0, w, w, out.ctypes.data) == 0
        assert np.array_equal(out, codes[y0:y0 + w, x0:x0 + w]), f"{label}: codes differ in window at ({x0},{y0})"


def _chase(codes, x, y):
    c = int(codes[y, x])
    if c & 2:
        return (x, y)
    if not c & 1:
        return None
    if (x & 1) != (y & 1):  # an edge takes the owner of the vertex or quad it is paired with
        d = (c >> 2) & 7
        x, y = x + DX[d], y + DY[d]
    while True:
        c = int(codes[y, x])
        if c & 2:
            return (x, y)
        if not c & 1:
            return None
        d = (c >> 2) & 7
        x, y = x + 2 * DX[d], y + 2 * DY[d]


def _cell_fn(win, x, y):
    f = np.float32
    i, j, dx, dy = x >> 1, y >> 1, x & 1, y & 1
    if not dx and not dy:
        return f(win[j, i])
    if dx and not dy:
        return f(f(f(0) + f(win[j, i + 1])) + f(win[j, i])) / f(2)
    if dy and not dx:
        return f(f(f(0) + f(win[j + 1, i])) + f(win[j, i])) / f(2)
    s = f(f(0) + f(win[j + 1, i + 1]))
    s = f(s + f(win[j, i + 1]))
    s = f(s + f(win[j, i]))
    s = f(s + f(win[j + 1, i]))
    return s / f(4)


def _check_tile(ds, win, rect, ext, seed, label, samples=4000, whole_grid=False):
    rng = np.random.default_rng(seed)
    codes = ds.codes()
    ch, cw = codes.shape
    crit_mask = (codes & 2) != 0
    K = ds.num_critical()
    assert K == int(np.count_nonzero(crit_mask)), f"{label}: K != number of critical flags"
    _involution(codes)
    _windows(codes, win, rect, ext, rng, label)
    _, _, own, crit = ds.arrays(flags=False, pairs=False, own=True, crit=True)
    ys, xs = np.nonzero(crit_mask)
    assert np.array_equal(crit, _pack(xs + ext[0], ys + ext[1])), f"{label}: critical list is not the raster order"
    del ys, xs, crit_mask
    fn, conn, _ = ds.critical_data()
    # owners of sampled cells against a sequential chase
    for _ in range(samples):
        x, y = int(rng.integers(0, cw)), int(rng.integers(0, ch))
        o = _chase(codes, x, y)
        exp = NONE if o is None else int(_pack(o[0] + ext[0], o[1] + ext[1]))
        assert int(own[y, x]) == exp, f"{label}: owner of ({x},{y})"
    # cell values and saddle incidence of sampled critical cells
    ks = rng.integers(0, K, size=min(K, samples))
    for k in ks:
        c = int(crit[k])
        x, y = (c & 0xFFFF) - ext[0], (c >> 16) - ext[1]
        assert np.float32(fn[k]).tobytes() == np.float32(_cell_fn(win, x, y)).tobytes(), f"{label}: fn of {k}"
        dx, dy = x & 1, y & 1
        if dx == dy:
            continue

        def idx(px, py):
            if px < 0 or py < 0 or px >= cw or py >= ch or own[py, px] == NONE:
                return NONE
            return int(np.searchsorted(crit, own[py, px]))
        assert [int(v) for v in conn[k]] == [idx(x + dx, y + dy), idx(x - dx, y - dy), idx(x + dy, y + dx),
                                             idx(x - dy, y - dx)], f"{label}: incidence of saddle {k}"
    if whole_grid:
        dims = (crit & 1) + ((crit >> 16) & 1)
        chi = int(np.count_nonzero(dims == 0)) - int(np.count_nonzero(dims == 1)) + int(np.count_nonzero(dims == 2))
        assert chi == 1, f"{label}: Euler characteristic {chi}"
    return K


def test_white_16384_whole_grid():
    data = mx.gen_white(SIDE, SIDE, seed=20241)
    print(f"[worst] white {SIDE}^2 whole grid", flush=True)
    rect = mx.grid_rect(SIDE, SIDE)
    buf = mx.DeviceBuffer.from_array(data)
    ds = mx.GridDataset(rect, rect)
    ds.bind_device(buf.ptr, SIDE)
    ds.work()
    K = _check_tile(ds, data, rect, rect, 1, "white16384", whole_grid=True)
    assert K > (2 * SIDE - 1) ** 2 // 10  # dense critical cells


@pytest.mark.parametrize("levels", [256, 0])
def test_16384_8x8_tiles(levels):
    data = mx.gen_white(SIDE, SIDE, levels=levels, seed=20242 + levels)
    print(f"[worst] white levels={levels} {SIDE}^2 8x8 tiles", flush=True)
    buf = mx.DeviceBuffer.from_array(data)
    pieces = split_pieces(SIDE, SIDE, lx=3, ly=3)
    assert len(pieces) == 64
    tiles = []
    for rect, ext in pieces:
        t = mx.GridDataset(rect, ext)
        t.bind_device(buf.ptr + 4 * ((ext[1] // 2) * SIDE + ext[0] // 2), SIDE)
        tiles.append(t)
    mx.work_tiles(tiles)
    for k in (0, 9, 27, 36, 54, 63):
        rect, ext = pieces[k]
        _check_tile(tiles[k], window(data, ext), rect, ext, 100 + k, f"levels={levels} leaf {k}", samples=1500)
    # full-size leaves against the oracle, word for word
    for k in (0, 27, 63):
        rect, ext = pieces[k]
        print(f"[worst] levels={levels} leaf {k} vs the oracle", flush=True)
        of, op, oo, oc = oracle.tile_work(window(data, ext), rect, ext)
        f, p, o, c = tiles[k].arrays()
        label = f"levels={levels} leaf {k}"
        assert np.array_equal(c, oc), f"{label}: critical list differs ({c.size} vs {oc.size})"
        assert np.array_equal(f, of), f"{label}: flags differ at {np.argwhere(f != of)[:5]}"
        assert np.array_equal(p, op), f"{label}: pairs differ at {np.argwhere(p != op)[:5]}"
        assert np.array_equal(o, oo), f"{label}: owners differ at {np.argwhere(o != oo)[:5]}"
        del of, op, oo, oc, f, p, o, c
