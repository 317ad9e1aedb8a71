"""GPU parity of the whole grid pipeline (mscx_grid_run: HIP stages for
every leaf tile, merge up, simplification, merge down, manifold discs)
against the oracle's GridDataManager order, word for word for every piece."""
import os

import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from tests.helpers import rng_field

pytestmark = pytest.mark.gpu


def _compare(data, label, **kw):
    print(f"[pipeline] {label} shape={data.shape} {kw}", flush=True)
    orc = oracle.Pipeline(data, keep=True, **kw)
    run = mx.GridRun(data, levels=kw.get("levels", 0), lx=kw.get("lx", 0), ly=kw.get("ly", 0),
                     threshold=kw.get("threshold", 0.0))
    assert run.num_pieces() == orc.num_pieces()
    for i in range(run.num_pieces()):
        a = orc.dump(4, i)
        b = run.piece(i).dump()
        assert a.size == b.size and np.array_equal(a, b), f"{label}: piece {i} differs"


@pytest.mark.parametrize("tiling", [dict(levels=0), dict(levels=1), dict(levels=3), dict(lx=1, ly=1),
                                    dict(lx=2, ly=2), dict(lx=3, ly=2)])
# Thresholds stay where the reference's simplification terminates: white
# noise at 0.1 already multiplies strangulated arcs without bound.
@pytest.mark.parametrize("threshold", [0.0, 0.01, 0.03])
def test_white(tiling, threshold):
    _compare(rng_field("white", 150, 130, 5), "white", threshold=threshold, **tiling)


@pytest.mark.parametrize("threshold", [0.0, 0.05])
def test_ties(threshold):
    _compare(rng_field("ties", 90, 70, 6), "ties", lx=1, ly=1, threshold=threshold)


def test_config_gaussians_512_l0():
    # the reference's CPU-runnable configuration: 512^2, 32 Gaussians, threshold 0
    _compare(mx.gen_gaussians(512, 512, 32, 1e-3, seed=1), "gauss512", levels=0, threshold=0.0)


@pytest.mark.timeout(300)
def test_white_512_single_leaf():
    """One leaf with 222,235 critical cells: the manifold pass splits a
    large leaf's label tables into chunks of 65,536 critical points, so this
    case reaches the later chunks (the 4096^2 bench leaves hold ~12k)."""
    _compare(mx.gen_white(512, 512, levels=0, seed=5), "white512-l0", levels=0, threshold=0.0)


def test_gaussians_512_slabs():
    _compare(mx.gen_gaussians(512, 512, 32, 1e-3, seed=1), "gauss512-l2", levels=2, threshold=0.01)


def test_value_noise_1024_4x4():
    _compare(mx.gen_value_noise(1024, 1024, period=256, seed=9), "vnoise1024", lx=2, ly=2, threshold=0.01)


def test_quantised_white_256():
    _compare(mx.gen_white(256, 256, levels=256, seed=3), "q256", lx=1, ly=1, threshold=0.005)


def test_dense_cancellation_fails_cleanly():
    # white noise at a 10% threshold: the reference grows arc multisets
    # without bound; the engine must return an error, not hang or OOM
    with pytest.raises(mx.MscxError) as e:
        mx.GridRun(rng_field("white", 150, 130, 5), levels=0, threshold=0.1)
    assert e.value.code == -4


@pytest.mark.timeout(600)
def test_bench_workload_4096_4x4():
    """The benchmark's workload itself: 4096^2 value noise (8 octaves,
    lacunarity 2, gain 0.5, period 1024, the bench seed), 4x4 tiles, 1%
    simplification; every one of the 31 pieces word for word after the
    discs (the oracle runs its --num-parallel form on up to 16 threads)."""
    g = mx.gen_value_noise(4096, 4096, 8, 2.0, 0.5, 1024.0, 20240601)
    _compare(g, "vnoise4096-bench", lx=2, ly=2, threshold=0.01, threads=min(16, os.cpu_count() or 1))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("threshold", [0.0, 0.005])
@pytest.mark.parametrize("levels", [0, 256])
def test_worst_case_fields_1024_8x8(levels, threshold):
    """The worst-case configuration's fields, tiling and both thresholds at a
    size the oracle finishes in seconds: white noise U[0,1) (dense critical
    cells) and its 256-level quantisation (ties on plateaus), 8x8 tiles,
    simplification at 0 and 0.5% of the range; every piece word for word
    after the discs."""
    g = mx.gen_white(1024, 1024, levels=levels, seed=77)
    _compare(g, f"white{levels}-1024-8x8", lx=3, ly=3, threshold=threshold, threads=min(16, os.cpu_count() or 1))


from tests.test_complex_cpu import DEGENERATE, TOO_FINE, oracle_outcome  # noqa: E402


@pytest.mark.parametrize("case", DEGENERATE, ids=[f"{a}x{b}-{kw}" for a, b, kw in DEGENERATE])
def test_tilings_near_the_grid_size(case):
    """Tilings as fine as the grid allows, and finer: where a piece's ext
    rect would leave the grid the run is refused (MSCX_EINVAL) before the
    device reads anything; otherwise every piece matches the oracle."""
    sx, sy, kw = case
    data = rng_field("white", sx, sy, 1)
    print(f"[pipeline] tiling {sx}x{sy} {kw}", flush=True)
    orc, msg = oracle_outcome(data, threshold=0.0, **kw)
    if msg is not None:
        assert msg.startswith(TOO_FINE), msg
        with pytest.raises(mx.MscxError) as e:
            mx.GridRun(data, threshold=0.0, **kw)
        assert e.value.code == -1 and TOO_FINE in str(e.value)
        return
    run = mx.GridRun(data, threshold=0.0, **kw)
    assert run.num_pieces() == orc.num_pieces()
    for i in range(run.num_pieces()):
        assert np.array_equal(orc.dump(4, i), run.piece(i).dump()), f"piece {i}"
