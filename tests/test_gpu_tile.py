"""GPU parity of the per-tile stages (gradient, collation, owners, saddle
incidence) against the oracle, bit-exact, through the C ABI.

Cases cover whole grids (the reference's L=0), slab tiles with halos
(createPieces_quadtree, grid_datamanager.cpp:53-89), 2-D tiles, ties,
signed zeros, constant fields, denormals and odd sizes around the kernels'
128-vertex tile width."""
import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from tests.helpers import rng_field, split_pieces, window

pytestmark = pytest.mark.gpu


def _check_tile(data, rect, ext, label):
    print(f"[tile] {label} rect={rect} ext={ext}", flush=True)
    win = window(data, ext)
    of, op, oo, oc = oracle.tile_work(win, rect, ext)
    ds = mx.GridDataset(rect, ext)
    ds.init(win)
    ds.work()
    f, p, o, c = ds.arrays()
    assert np.array_equal(c, oc), f"{label}: critical list differs ({c.size} vs {oc.size})"
    assert np.array_equal(f, of), f"{label}: flags differ at {np.argwhere(f != of)[:5]}"
    assert np.array_equal(p, op), f"{label}: pairs differ at {np.argwhere(p != op)[:5]}"
    assert np.array_equal(o, oo), f"{label}: owners differ at {np.argwhere(o != oo)[:5]}"
    # per-critical data vs the oracle's writeout_connectivity
    _, cps = oracle.parse_dump(oracle.tile_complex(win, rect, ext))
    fn, conn, pr = ds.critical_data()
    fnb = fn.view(np.uint32)
    assert len(cps) == c.size
    for k, cp in enumerate(cps):
        assert cp["cell"] == int(c[k])
        assert cp["fn"] == int(fnb[k]), f"{label}: fn bits differ at cp {k}"
        assert cp["pair"] == int(pr[k]), f"{label}: pair idx differs at cp {k}"
        x, y = mx.unpack(c[k])
        if (x & 1) != (y & 1):
            des = sorted(int(v) for v in conn[k, :2] if v != 0xFFFFFFFF)
            asc = sorted(int(v) for v in conn[k, 2:] if v != 0xFFFFFFFF)
            assert des == sorted(cp["des"]), f"{label}: saddle {k} des differs"
            assert asc == sorted(cp["asc"]), f"{label}: saddle {k} asc differs"


def test_h1_witness():
    # CPU path pairs edge (1,2) with quad (1,3) (ascending compareCells,
    # grid_dataset.cpp:1530-1535); the OpenCL path would pick (1,1).
    data = np.array([[2, 3], [5, 6], [4, 1]], np.float32)
    ds = mx.GridDataset(mx.grid_rect(2, 3))
    ds.init(data)
    ds.work()
    f, p, o, c = ds.arrays()
    assert mx.unpack(p[2, 1]) == (1, 3)
    assert [mx.unpack(v) for v in c] == [(0, 0), (1, 1), (2, 1), (0, 3), (2, 4)]
    _check_tile(data, mx.grid_rect(2, 3), mx.grid_rect(2, 3), "h1")


@pytest.mark.parametrize("kind", ["white", "ties", "const", "signedzero", "smooth", "denormal"])
@pytest.mark.parametrize("shape", [(2, 2), (3, 3), (2, 7), (7, 2), (5, 9), (33, 17), (130, 9), (9, 131), (257, 66)])
def test_whole_grid(kind, shape):
    sx, sy = shape
    data = rng_field(kind, sx, sy, seed=sx * 1000 + sy)
    _check_tile(data, mx.grid_rect(sx, sy), mx.grid_rect(sx, sy), f"{kind}{shape}")


@pytest.mark.parametrize("kind", ["white", "ties", "smooth"])
@pytest.mark.parametrize("tiling", [dict(levels=1), dict(levels=2), dict(levels=3), dict(lx=1, ly=1),
                                    dict(lx=2, ly=2), dict(lx=2, ly=1)])
def test_tiles_with_halos(kind, tiling):
    sx, sy = 61, 47
    data = rng_field(kind, sx, sy, seed=7)
    for i, (rect, ext) in enumerate(split_pieces(sx, sy, **tiling)):
        _check_tile(data, rect, ext, f"{kind}{tiling}#{i}")


def test_batched_tiles_match_single():
    sx, sy = 200, 150
    data = rng_field("white", sx, sy, seed=3)
    pieces = split_pieces(sx, sy, lx=2, ly=2)
    tiles = []
    for rect, ext in pieces:
        t = mx.GridDataset(rect, ext)
        t.init(window(data, ext))
        tiles.append(t)
    mx.work_tiles(tiles)
    for (rect, ext), t in zip(pieces, tiles):
        of, op, oo, oc = oracle.tile_work(window(data, ext), rect, ext)
        f, p, o, c = t.arrays()
        assert np.array_equal(c, oc) and np.array_equal(f, of) and np.array_equal(p, op) and np.array_equal(o, oo)


def test_config_gaussians_512():
    data = mx.gen_gaussians(512, 512, 32, 1e-3, seed=1)
    _check_tile(data, mx.grid_rect(512, 512), mx.grid_rect(512, 512), "gauss512")


def test_value_noise_tiles():
    data = mx.gen_value_noise(300, 300, seed=5, period=60)
    for i, (rect, ext) in enumerate(split_pieces(300, 300, lx=2, ly=2)):
        _check_tile(data, rect, ext, f"vnoise#{i}")
