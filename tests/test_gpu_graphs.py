"""The tile stages that need no host decision (descriptor upload, gradient,
scan, ring resolution, owner images, count download) submitted as one HIP
graph, built once per launch configuration and replayed.  Every replay,
on new data written in place into the bound device buffer, must equal the
directly launched run of the same data word for word; direct runs and host
synchronisations sit between the replays, and a second configuration is
built and replayed in turn."""
import numpy as np
import pytest

import mscomplex_amd as mx
from tests.helpers import split_pieces

pytestmark = pytest.mark.gpu


def _bound_tiles(buf, sx, sy, lx, ly):
    tiles = []
    for rect, ext in split_pieces(sx, sy, lx=lx, ly=ly):
        t = mx.GridDataset(rect, ext)
        t.bind_device(buf.ptr + 4 * ((ext[1] // 2) * sx + ext[0] // 2), sx)
        tiles.append(t)
    return tiles


def _snapshot(tiles):
    out = []
    for t in tiles:
        _, _, own, crit = t.arrays(flags=False, pairs=False, own=True, crit=True)
        fn, conn, pr = t.critical_data()
        out.append((t.codes(), own, crit, fn.view(np.uint32), conn, pr))
    return out


def _same(a, b, label):
    assert len(a) == len(b), label
    for k, (x, y) in enumerate(zip(a, b)):
        for i, (u, v) in enumerate(zip(x, y)):
            assert np.array_equal(u, v), f"{label}: tile {k} array {i}"


def _fields(sx, sy, seed, n):
    return [mx.gen_value_noise(sx, sy, seed=seed + i, period=max(sx, sy) / 3) if i % 2 else
            mx.gen_gaussians(sx, sy, 16, 1e-3, seed=seed + i) for i in range(n)]


@pytest.mark.parametrize("shape", [(300, 220, 1, 1), (513, 190, 2, 1)])
def test_replays_equal_direct_runs(shape):
    sx, sy, lx, ly = shape
    print(f"[graphs] {sx}x{sy} tiles {lx},{ly}", flush=True)
    fields = _fields(sx, sy, 100 + sx, 4)
    gbuf, ebuf = mx.DeviceBuffer(4 * sx * sy), mx.DeviceBuffer(4 * sx * sy)
    gt, et = _bound_tiles(gbuf, sx, sy, lx, ly), _bound_tiles(ebuf, sx, sy, lx, ly)
    prev = mx.set_graphs(True)
    try:
        c0, r0 = mx.graph_stats()
        for k, f in enumerate(fields):
            gbuf.upload(f)
            mx.work_tiles(gt)                 # built on the first pass, replayed after
            got = _snapshot(gt)
            mx.set_graphs(False)
            ebuf.upload(f)
            mx.work_tiles(et)                 # the same data, launched directly
            want = _snapshot(et)
            mx.set_graphs(True)
            _same(got, want, f"{sx}x{sy} field {k}")
        c1, r1 = mx.graph_stats()
        assert c1 - c0 == 1 and r1 - r0 == len(fields) - 1, (c0, r0, c1, r1)
    finally:
        mx.set_graphs(prev)


def test_pipeline_with_graphs_equals_without():
    g = mx.gen_value_noise(400, 330, seed=9, period=120)
    prev = mx.set_graphs(True)
    try:
        runs = [mx.GridRun(g, lx=1, ly=1, threshold=0.02) for _ in range(3)]
        mx.set_graphs(False)
        ref = mx.GridRun(g, lx=1, ly=1, threshold=0.02)
    finally:
        mx.set_graphs(prev)
    for r in runs:
        assert r.num_pieces() == ref.num_pieces()
        for p in range(ref.num_pieces()):
            assert np.array_equal(r.piece(p).dump(), ref.piece(p).dump()), f"piece {p}"
