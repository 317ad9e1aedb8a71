"""The math-fixed pins of the merge tree, cancellation, simplification and
discs (tests/pins.py, checked on the oracle in test_oracle_pins_cpu.py) on
the HIP pipeline's own output, at the configurations' fields and tilings:

  (i)   the root merged from GPU-worked leaf tiles (mscx_complex_merge_up at
        threshold 0), canonicalised by cell id, equals the complex of one
        whole-grid tile;
  (ii)  the root's survivors after simplification form the Morse complex of
        a disc over Z/2;
  (iii) every leaf's surviving extrema's discs partition its rect;
  (iv)  every leaf's surviving nodes are the root's survivors in its rect.

Word-for-word parity with the oracle at these sizes is in
test_gpu_pipeline.py; these pins also run where no oracle does (the 16384^2
worst case, test_gpu_worstcase_e2e.py)."""
import numpy as np
import pytest

import mscomplex_amd as mx
from tests import pins
from tests.helpers import split_pieces, window

pytestmark = pytest.mark.gpu


def merged_root(data, **tiling):
    """GPU work() on every leaf tile, writeout_connectivity, then merge_up
    pairwise in the GridDataManager order (level by level, (j, j+1))."""
    sy, sx = data.shape
    level = []
    for rect, ext in split_pieces(sx, sy, **tiling):
        t = mx.GridDataset(rect, ext)
        t.init(window(data, ext))
        t.work()
        level.append(mx.GridMSComplex.from_dataset(t))
    while len(level) > 1:
        level = [mx.GridMSComplex.merge_up(level[j], level[j + 1]) for j in range(0, len(level), 2)]
    return level[0]


def whole_complex(data):
    sy, sx = data.shape
    t = mx.GridDataset(mx.grid_rect(sx, sy))
    t.init(data)
    t.work()
    return mx.GridMSComplex.from_dataset(t)


def check_run(run, label):
    """(ii)-(iv) on a finished GridRun."""
    npc = run.num_pieces()
    rd = run.piece(npc - 1).dump()
    rn = pins.nodes(rd)
    rmask = (rn["flags"] & pins.F_BCANC) == 0
    h = pins.assert_disc_homology(rd, rmask, (label, "root"))
    rs = rn["cell"][rmask]
    nleaves = (npc + 1) // 2
    for leaf in range(nleaves):
        d = run.piece(leaf).dump()
        n = pins.nodes(d)
        rect = n["rect"]
        pins.assert_disc_partition(d, rect, (label, leaf))
        ls = n["cell"][((n["flags"] & pins.F_BCANC) == 0) & pins.cells_in(n["cell"], rect)]
        assert np.array_equal(np.sort(ls), np.sort(rs[pins.cells_in(rs, rect)])), (label, leaf)
    return h


CASES = [
    ("gauss512", lambda: mx.gen_gaussians(512, 512, 32, 1e-3, seed=1), dict(lx=1, ly=1), [0.0]),
    ("vnoise1024-4x4", lambda: mx.gen_value_noise(1024, 1024, period=256, seed=9), dict(lx=2, ly=2), [0.0, 0.01]),
    ("white1024-8x8", lambda: mx.gen_white(1024, 1024, levels=0, seed=77), dict(lx=3, ly=3), [0.0, 0.005]),
    ("q256-1024-8x8", lambda: mx.gen_white(1024, 1024, levels=256, seed=77), dict(lx=3, ly=3), [0.0, 0.005]),
]


@pytest.mark.timeout(300)
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_merged_root_equals_whole_grid(case):
    label, gen, tiling, _ = case
    data = gen()
    print(f"[pins] (i) {label} {tiling}", flush=True)
    a = pins.canon(merged_root(data, **tiling).dump())
    b = pins.canon(whole_complex(data).dump())
    assert np.array_equal(a, b), label


@pytest.mark.timeout(300)
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_pipeline_pins(case):
    label, gen, tiling, thresholds = case
    data = gen()
    for th in thresholds:
        print(f"[pins] (ii)-(iv) {label} {tiling} threshold={th}", flush=True)
        h = check_run(mx.GridRun(data, threshold=th, **tiling), f"{label}@{th}")
        assert h["saddle"] > 0


def test_discs_are_traced_on_the_gpu():
    """The saddle 1-manifolds of the manifold pass come from the GPU tracing
    (k_disc_trace), not from the host fallback, and equal the oracle's."""
    import oracle
    g = mx.gen_value_noise(600, 520, seed=13, period=150)
    before = mx.disc_trace_stats()
    run = mx.GridRun(g, lx=1, ly=1, threshold=0.01)
    after = mx.disc_trace_stats()
    gpu, host, dropped, wide = (a - b for a, b in zip(after[:4], before[:4]))
    print(f"[pins] disc slots: {gpu} on the GPU ({wide} by a wavefront), {host} on the host, {dropped} calls "
          f"dropped; longest list {after[4]}", flush=True)
    assert dropped == 0 and gpu > 0 and gpu > 10 * host
    assert wide > 0, "no slot took the wavefront tracing"
    orc = oracle.Pipeline(g, lx=1, ly=1, threshold=0.01)
    for i in range(run.num_pieces()):
        assert np.array_equal(run.piece(i).dump(), orc.dump(4, i)), i


def test_discs_traced_by_wavefronts_and_the_host():
    """White noise at 3%: some saddle discs grow more chains than a lane
    follows (a wavefront traces them) and a few more than a wavefront's
    chain table holds (the host traces those); every list equals the
    oracle's."""
    import oracle
    from tests.helpers import rng_field
    g = rng_field("white", 300, 300, 3)
    print("[pins] white 300^2 at 3%: wavefront and host disc slots", flush=True)
    before = mx.disc_trace_stats()
    run = mx.GridRun(g, levels=0, threshold=0.03)
    after = mx.disc_trace_stats()
    gpu, host, dropped, wide = (a - b for a, b in zip(after[:4], before[:4]))
    print(f"[pins] disc slots: {gpu} on the GPU ({wide} by a wavefront), {host} on the host", flush=True)
    assert dropped == 0 and wide > 0 and host > 0
    orc = oracle.Pipeline(g, levels=0, threshold=0.03)
    assert np.array_equal(run.piece(0).dump(), orc.dump(4, 0))
