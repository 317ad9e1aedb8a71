"""Pins of the oracle (the CPU restatement of the reference's grid path)
that do not depend on re-typing its formulas: a hand-derived witness,
invariants, brute-force recomputation of owners and collation, closed forms
of cell values, and special fields with known topology."""
import numpy as np
import pytest

import oracle
from tests.helpers import rng_field, split_pieces, window


def unpack(p):
    p = int(p)
    return (p & 0xFFFF, p >> 16) if p != 0xFFFFFFFF else (-1, -1)


def whole(sx, sy):
    return (0, 0, 2 * (sx - 1), 2 * (sy - 1))


def test_h1_witness():
    # On this 2x3 grid the reference's CPU path pairs edge (1,2) with quad
    # (1,3) (compareCells sorts ascending, grid_dataset.cpp); the descending
    # compare of its OpenCL path would pick (1,1).
    data = np.array([[2, 3], [5, 6], [4, 1]], np.float32)
    f, p, o, c = oracle.tile_work(data, whole(2, 3))
    assert unpack(p[2, 1]) == (1, 3)
    assert [unpack(v) for v in c] == [(0, 0), (1, 1), (2, 1), (0, 3), (2, 4)]


def _dim(x, y):
    return (x & 1) + (y & 1)


@pytest.mark.parametrize("kind", ["white", "ties", "const", "smooth", "signedzero"])
@pytest.mark.parametrize("shape", [(3, 3), (9, 7), (31, 17)])
def test_invariants_whole_grid(kind, shape):
    sx, sy = shape
    data = rng_field(kind, sx, sy, seed=sx + 7 * sy)
    f, p, o, c = oracle.tile_work(data, whole(sx, sy))
    ch, cw = f.shape
    nd = [0, 0, 0]
    for y in range(ch):
        for x in range(cw):
            fl = int(f[y, x])
            assert fl in (1, 2), "L=0: every cell is paired or critical, never both"
            if fl == 1:
                px, py = unpack(p[y, x])
                assert abs(px - x) + abs(py - y) == 1
                assert abs(_dim(px, py) - _dim(x, y)) == 1
                assert unpack(p[py, px]) == (x, y), "pairing is an involution"
            else:
                nd[_dim(x, y)] += 1
    assert nd[0] - nd[1] + nd[2] == 1, "Euler characteristic of a disc"
    # collation = every critical cell in raster order
    exp = [(x, y) for y in range(ch) for x in range(cw) if f[y, x] & 2]
    assert [unpack(v) for v in c] == exp


def _chase_owner(f, p, x, y, ext):
    """Follow the gradient from a vertex (down to a minimum) or a quad (up to
    a maximum) by brute force; local coordinates."""
    while not (f[y, x] & 2):
        ex, ey = unpack(p[y, x])
        ex, ey = ex - ext[0], ey - ext[1]
        x, y = 2 * ex - x, 2 * ey - y  # across the paired edge
    return (x, y)


@pytest.mark.parametrize("kind", ["white", "ties", "smooth"])
@pytest.mark.parametrize("tiling", [dict(levels=0), dict(levels=2), dict(lx=1, ly=1)])
def test_owners_brute_force(kind, tiling):
    sx, sy = 25, 21
    data = rng_field(kind, sx, sy, seed=3)
    for rect, ext in split_pieces(sx, sy, **tiling):
        f, p, o, c = oracle.tile_work(window(data, ext), rect, ext)
        ch, cw = f.shape
        for y in range(ch):
            for x in range(cw):
                gx, gy = x + ext[0], y + ext[1]
                inr = rect[0] <= gx <= rect[2] and rect[1] <= gy <= rect[3]
                d = _dim(gx, gy)
                if f[y, x] & 2:
                    exp = (gx, gy)
                elif not inr:
                    exp = (-1, -1)
                elif d != 1:
                    ox, oy = _chase_owner(f, p, x, y, ext)
                    exp = (ox + ext[0], oy + ext[1])
                else:  # edge: owner of its paired vertex / quad
                    px, py = unpack(p[y, x])
                    exp = unpack(o[py - ext[1], px - ext[0]])
                assert unpack(o[y, x]) == exp, (rect, ext, gx, gy)


def test_cell_values_closed_form():
    data = rng_field("white", 13, 11, seed=5)
    buf = oracle.tile_complex(data, whole(13, 11))
    _, cps = oracle.parse_dump(buf)
    for cp in cps:
        x, y = unpack(cp["cell"])
        i, j = x // 2, y // 2
        if x % 2 == 0 and y % 2 == 0:
            pts = [(i, j)]
        elif x % 2 == 1 and y % 2 == 0:
            pts = [(i + 1, j), (i, j)]
        elif x % 2 == 0:
            pts = [(i, j + 1), (i, j)]
        else:
            pts = [(i + 1, j + 1), (i + 1, j), (i, j), (i, j + 1)]
        acc = np.float32(0.0)
        for a, b in pts:
            acc = np.float32(acc + data[b, a])
        val = np.float32(acc / np.float32(len(pts)))
        assert np.array_equal(np.array([val]).view(np.uint32), np.array([cp["fn"]], np.uint32))


def test_monotone_field_single_minimum():
    sx, sy = 20, 15
    y, x = np.mgrid[0:sy, 0:sx]
    data = (x + 1e-3 * y).astype(np.float32)
    f, p, o, c = oracle.tile_work(data, whole(sx, sy))
    dims = [_dim(*unpack(v)) for v in c]
    assert dims.count(0) == 1 and unpack(c[0]) == (0, 0)
    assert dims.count(0) - dims.count(1) + dims.count(2) == 1


def test_radial_bump_single_interior_maximum():
    sx = sy = 21
    y, x = np.mgrid[0:sy, 0:sx]
    data = np.exp(-((x - 10.3) ** 2 + (y - 9.7) ** 2) / 30.0).astype(np.float32)
    f, p, o, c = oracle.tile_work(data, whole(sx, sy))
    maxima = [unpack(v) for v in c if _dim(*unpack(v)) == 2]
    assert len(maxima) == 1
    mx_, my_ = maxima[0]
    assert 16 <= mx_ <= 24 and 16 <= my_ <= 24


def test_threshold_zero_cancels_nothing_and_merges_keep_euler():
    data = rng_field("white", 33, 27, seed=1)
    pl = oracle.Pipeline(data, levels=2, threshold=0.0)
    for piece in range(pl.num_pieces()):
        _, cps = oracle.parse_dump(pl.dump(1, piece))
        live = [cp for cp in cps if not cp["flags"] & 1]
        dims = [_dim(*unpack(cp["cell"])) for cp in live]
        if piece == pl.num_pieces() - 1:
            assert dims.count(0) - dims.count(1) + dims.count(2) == 1
    a = pl.dump(1, pl.num_pieces() - 1)
    b = pl.dump(2, pl.num_pieces() - 1)
    assert np.array_equal(a, b), "simplify at threshold 0 must not change the root"


def test_simplify_cancels_adjacent_dimensions_and_keeps_euler():
    data = rng_field("smooth", 40, 30, seed=2)
    pl = oracle.Pipeline(data, levels=0, threshold=0.05)
    _, before = oracle.parse_dump(pl.dump(1, 0))
    _, after = oracle.parse_dump(pl.dump(2, 0))
    newly = [i for i, cp in enumerate(after) if cp["flags"] & 4 and not before[i]["flags"] & 4]
    assert newly, "expected cancellations at 5%"
    live = [i for i in range(len(after)) if i not in set(newly)]
    for i in newly:
        j = after[i]["pair"]
        assert after[j]["pair"] == i
        assert abs(_dim(*unpack(after[i]["cell"])) - _dim(*unpack(after[j]["cell"]))) == 1
    dims = [_dim(*unpack(after[i]["cell"])) for i in live]
    assert dims.count(0) - dims.count(1) + dims.count(2) == 1


@pytest.mark.parametrize("levels,lx,ly", [(3, 0, 0), (0, 2, 2), (0, 2, 1)])
def test_parallel_pipeline_equals_single_threaded(levels, lx, ly):
    """The oracle's --num-parallel form (independent tiles and merge-tree
    nodes of one level on several threads) gives the same complexes after
    every phase as the -st order."""
    f = rng_field("smooth", 200, 170, seed=12)
    a = oracle.Pipeline(f, levels=levels, lx=lx, ly=ly, threshold=0.02)
    b = oracle.Pipeline(f, levels=levels, lx=lx, ly=ly, threshold=0.02, threads=6)
    assert a.num_pieces() == b.num_pieces()
    for st in range(5):
        for p in range(a.num_pieces()):
            assert np.array_equal(a.dump(st, p), b.dump(st, p)), (st, p)
