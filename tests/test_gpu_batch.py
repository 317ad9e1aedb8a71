"""Batched runs (mscx_batch_run): many uneven grids share every stage
launch; each grid's complexes must equal the oracle's and the single-grid
run's, word for word, and a failing grid must not disturb the others."""
import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from tests.helpers import rng_field

pytestmark = pytest.mark.gpu


def _mix(n, seed):
    r = np.random.default_rng(seed)
    grids, lx, ly, th = [], [], [], []
    for i in range(n):
        sx = int(np.exp(r.uniform(np.log(16), np.log(200))))
        sy = max(2, int(sx * r.uniform(0.25, 1.0))) if r.random() < 0.5 else int(sx / r.uniform(0.25, 1.0))
        sy = min(max(sy, 2), 400)
        if i % 2:
            g = mx.gen_gaussians(sx, sy, 8, 1e-3, seed=seed * 1000 + i)
        else:
            g = mx.gen_value_noise(sx, sy, seed=seed * 1000 + i, period=max(sx, sy) / 4)
        grids.append(g)
        lx.append(int(r.integers(0, 3)) if min(sx, sy) > 40 else 0)
        ly.append(int(r.integers(0, 3)) if min(sx, sy) > 40 else 0)
        th.append(float(r.choice([0.0, 0.005, 0.01, 0.02])))
    return grids, lx, ly, th


def test_batch_matches_oracle_and_single_runs():
    grids, lx, ly, th = _mix(24, seed=1)
    print(f"[batch] {len(grids)} grids", flush=True)
    b = mx.Batch(grids, lx, ly, th)
    for i, g in enumerate(grids):
        view = b.grid(i)
        kw = dict(lx=lx[i], ly=ly[i], threshold=th[i]) if (lx[i] or ly[i]) else dict(levels=0, threshold=th[i])
        ref = oracle.Pipeline(g, **kw)
        single = mx.GridRun(g, levels=0, lx=lx[i], ly=ly[i], threshold=th[i])
        assert view.num_pieces() == ref.num_pieces() == single.num_pieces()
        for p in range(view.num_pieces()):
            a, c, d = ref.dump(4, p), view.piece(p).dump(), single.piece(p).dump()
            assert np.array_equal(a, c), f"grid {i} piece {p}: batch differs from the oracle"
            assert np.array_equal(c, d), f"grid {i} piece {p}: batch differs from the single run"


def test_failing_grid_is_isolated():
    good = [mx.gen_value_noise(64, 48, seed=3), mx.gen_gaussians(50, 70, 8, 1e-3, seed=4)]
    bad_nan = np.full((20, 20), np.nan, np.float32)
    explode = rng_field("white", 150, 130, 5)  # 10% threshold: arc explosion
    b = mx.Batch([good[0], bad_nan, explode, good[1]], [0, 0, 0, 0], [0, 0, 0, 0], [0.01, 0.01, 0.1, 0.01])
    assert b.status(1)[0] == -1 and "NaN" in b.status(1)[1]
    assert b.status(2)[0] == -4
    for i, g in ((0, good[0]), (3, good[1])):
        ref = oracle.Pipeline(g, levels=0, threshold=0.01)
        assert np.array_equal(ref.dump(4, 0), b.grid(i).piece(0).dump())


def test_batch_at_scale_equals_plain_runs():
    # the batched configuration's field mix at a larger scale than the oracle
    # reaches in a test: 96 grids with long sides log-uniform in [64, 1024],
    # aspect up to 1:4, value noise at 1% and Gaussian sums at 0, various
    # tilings; every grid's pieces equal the plain run's word for word
    # (checksums of the dump words)
    mix = mx.uneven_mix(96, 64, 1024, 5, aspect_max=4.0, thresholds=0.01)
    grids = [g for g, _ in mix]
    th = [0.0 if i % 2 else t for i, (_, t) in enumerate(mix)]
    r = np.random.default_rng(5)
    lx = [int(r.integers(0, 3)) if min(g.shape) > 200 else 0 for g in grids]
    ly = [int(r.integers(0, 3)) if min(g.shape) > 200 else 0 for g in grids]
    print(f"[batch] {len(grids)} grids, {sum(g.size for g in grids)} vertices", flush=True)
    b = mx.Batch(grids, lx, ly, th)
    for i, g in enumerate(grids):
        assert b.status(i)[0] == 0, b.status(i)
        single = mx.GridRun(g, levels=0, lx=lx[i], ly=ly[i], threshold=th[i])
        assert b.grid(i).checksum() == single.checksum(), f"grid {i} {g.shape} differs from its plain run"
