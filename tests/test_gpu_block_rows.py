"""Kernel block tile height (mscx_set_block_rows): tiles whose gradient,
trace and owner kernels run on 16-row blocks give the same results, word for
word, as the default 32-row blocks and as the oracle -- on tile heights
around multiples of both heights, on boundary and tie fields, in one batch
that mixes both heights, and through the no-allocation entry point."""
import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from tests.helpers import rng_field, split_pieces, window

pytestmark = pytest.mark.gpu


@pytest.fixture
def rows16():
    mx.set_block_rows([(1 << 30, 16)])
    yield
    mx.set_block_rows(None)


def _dumps(run):
    return [run.piece(i).dump() for i in range(run.num_pieces())]


@pytest.mark.parametrize("shape", [(15, 70), (16, 33), (17, 64), (31, 65), (33, 96), (48, 130), (97, 129)])
@pytest.mark.parametrize("kind", ["white", "ties"])
def test_tile_stages_match_the_oracle(rows16, shape, kind):
    sy, sx = shape
    data = rng_field(kind, sx, sy, sy * 131 + sx)
    print(f"[rows16] {kind} {shape}", flush=True)
    rect = ext = (0, 0, 2 * (sx - 1), 2 * (sy - 1))
    t = mx.GridDataset(rect, ext)
    t.init(data)
    t.work()
    flags, pairs, own, crit = t.arrays()
    of, op, oo, oc = oracle.tile_work(data, rect, ext)
    assert np.array_equal(crit, oc)
    assert np.array_equal(flags, of) and np.array_equal(own, oo) and np.array_equal(pairs, op)


@pytest.mark.parametrize("tiling", [dict(levels=2), dict(levels=0, lx=1, ly=2)])
def test_pipeline_equals_default_height(tiling):
    data = mx.gen_value_noise(300, 260, seed=4, period=64)
    print(f"[rows16] pipeline {tiling}", flush=True)
    ref = _dumps(mx.GridRun(data, threshold=0.01, **tiling))
    mx.set_block_rows([(1 << 30, 16)])
    try:
        got = _dumps(mx.GridRun(data, threshold=0.01, **tiling))
    finally:
        mx.set_block_rows(None)
    assert len(ref) == len(got)
    for i, (a, b) in enumerate(zip(ref, got)):
        assert a.size == b.size and np.array_equal(a, b), f"piece {i} differs"
    orc = oracle.Pipeline(data, threshold=0.01, keep=True, **tiling)
    for i, b in enumerate(got):
        a = orc.dump(4, i)
        assert a.size == b.size and np.array_equal(a, b), f"piece {i} differs from the oracle"


def test_batch_mixing_both_heights():
    grids = [rng_field("white", 40 + 13 * k, 30 + 29 * k, k) for k in range(8)]
    ths = [0.01] * len(grids)
    print("[rows16] batch", flush=True)
    plain = [_dumps(mx.GridRun(g, levels=0, threshold=t)) for g, t in zip(grids, ths)]
    mx.set_block_rows([(120, 16)])  # small tiles on 16-row blocks, the rest on 32
    try:
        b = mx.Batch(grids, 0, 0, ths)
        for i in range(len(grids)):
            assert b.status(i)[0] == 0
            got = _dumps(b.grid(i))
            assert [x.tobytes() for x in got] == [x.tobytes() for x in plain[i]], f"grid {i} differs"
    finally:
        mx.set_block_rows(None)


def test_noalloc_entry_on_16_row_blocks(rows16):
    from tests.test_zz_no_alloc import _noalloc, _reference, _same
    side = 512
    data = mx.gen_value_noise(side, side, seed=8, period=128)
    dev = mx.DeviceBuffer.from_array(data)
    for rect, ext in (split_pieces(side, side, lx=1, ly=1)[2], split_pieces(side, side, lx=2, ly=2)[9]):
        print(f"[rows16] noalloc {rect}", flush=True)
        ref = _reference(data, rect, ext)
        rc, k, got = _noalloc(dev, side, rect, ext, ref[0].size)
        assert rc == 0 and k == ref[0].size
        _same(got, ref)
