"""The merge tree on the GPU (MSCX_RUN_DEVICE_MERGE): every piece of the
device-merged run equals the oracle's GridDataManager order word for word
(merge up, simplification, merge down, discs), and the device path really
ran (no fallback to the host merges)."""
import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from tests.helpers import rng_field

pytestmark = pytest.mark.gpu


def _run_device(data, label, **kw):
    print(f"[dmerge] {label} shape={data.shape} {kw}", flush=True)
    s0 = mx.merge_stats()
    run = mx.GridRun(data, levels=kw.get("levels", 0), lx=kw.get("lx", 0), ly=kw.get("ly", 0),
                     threshold=kw.get("threshold", 0.0), merge="device")
    s1 = mx.merge_stats()
    nmerge = run.num_pieces() - (run.num_pieces() + 1) // 2
    assert s1[1] == s0[1], f"{label}: the device tree fell back to the host path"
    assert s1[0] - s0[0] == 2 * nmerge, f"{label}: {s1[0] - s0[0]} device merges for {nmerge} merge nodes"
    return run


def _compare_oracle(data, label, **kw):
    run = _run_device(data, label, **kw)
    orc = oracle.Pipeline(data, keep=True, **kw)
    assert run.num_pieces() == orc.num_pieces()
    for i in range(run.num_pieces()):
        a = orc.dump(4, i)
        b = run.piece(i).dump()
        assert a.size == b.size and np.array_equal(a, b), f"{label}: piece {i} differs from the oracle"


@pytest.mark.parametrize("tiling", [dict(levels=1), dict(levels=3), dict(lx=1, ly=1), dict(lx=2, ly=2),
                                    dict(lx=3, ly=2)])
@pytest.mark.parametrize("threshold", [0.0, 0.01, 0.03])
def test_white(tiling, threshold):
    _compare_oracle(rng_field("white", 150, 130, 5), "white", threshold=threshold, **tiling)


@pytest.mark.parametrize("threshold", [0.0, 0.05])
def test_ties(threshold):
    _compare_oracle(rng_field("ties", 90, 70, 6), "ties", lx=1, ly=1, threshold=threshold)


def test_gaussians_512_4x4():
    _compare_oracle(mx.gen_gaussians(512, 512, 32, 1e-3, seed=1), "gauss512", lx=2, ly=2, threshold=0.0)


def test_value_noise_1024_4x4():
    d = mx.gen_value_noise(1024, 1024, 8, 2.0, 0.5, 256.0, 9)
    _compare_oracle(d, "vnoise1024", lx=2, ly=2, threshold=0.01)


@pytest.mark.parametrize("levels,threshold", [(0, 0.0), (256, 0.005)])
def test_worst_case_fields_1024_8x8(levels, threshold):
    _compare_oracle(mx.gen_white(1024, 1024, levels=levels, seed=11), f"white1024-q{levels}", lx=3, ly=3,
                    threshold=threshold)


def test_bench_field_2048_device_equals_host():
    # the benchmark's field family at 2048^2, 4x4 tiles, 1%: device tree vs
    # the host tree, every piece
    d = mx.gen_value_noise(2048, 2048, 8, 2.0, 0.5, 512.0, 20240601)
    dev = _run_device(d, "vnoise2048", lx=2, ly=2, threshold=0.01)
    host = mx.GridRun(d, lx=2, ly=2, threshold=0.01, merge="host")
    assert dev.num_pieces() == host.num_pieces()
    for i in range(dev.num_pieces()):
        a, b = host.piece(i).dump(), dev.piece(i).dump()
        assert a.size == b.size and np.array_equal(a, b), f"piece {i} differs between host and device trees"


def test_single_leaf_and_no_discs():
    d = rng_field("white", 64, 48, 3)
    for tiling in (dict(levels=0), dict(lx=1, ly=0)):
        run = mx.GridRun(d, threshold=0.02, merge="device", discs=False, **tiling)
        ref = oracle.Pipeline(d, threshold=0.02, keep=True, **tiling)
        for i in range(run.num_pieces()):
            # no discs were asked for: the complexes after merge-down
            a, b = ref.dump(3, i), run.piece(i).dump()
            assert a.size == b.size and np.array_equal(a, b), f"{tiling}: piece {i} differs"
