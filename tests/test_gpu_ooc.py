"""Out-of-core runs (MSCX_RUN_OUT_OF_CORE, the reference's --out-of-core,
grid_datamanager.cpp:91-119): pieces that no phase is using are kept
serialised -- as archive words in host memory, or as archive files under
MSCX_SPILL_DIR -- and every piece of the run equals the in-core run word for
word (and the oracle's), also from a file with disc files written."""
import os

import numpy as np
import pytest

import mscomplex_amd as mx
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spill_dir", [False, True])
@pytest.mark.parametrize("tiling,threshold", [(dict(lx=2, ly=2), 0.01), (dict(levels=3), 0.0), (dict(lx=1, ly=1), 0.02)])
def test_out_of_core_equals_in_core(tmp_path, monkeypatch, spill_dir, tiling, threshold):
    if spill_dir:
        monkeypatch.setenv("MSCX_SPILL_DIR", str(tmp_path))
    g = mx.gen_value_noise(300, 260, seed=4, period=64)
    print(f"[ooc] {tiling} th={threshold} dir={spill_dir}", flush=True)
    a = mx.GridRun(g, threshold=threshold, **tiling)
    b = mx.GridRun(g, threshold=threshold, out_of_core=True, **tiling)
    orc = oracle.Pipeline(g, threshold=threshold, **tiling)
    assert a.num_pieces() == b.num_pieces() == orc.num_pieces()
    # the pieces come back in an order other than the one they were spilled in
    for i in reversed(range(b.num_pieces())):
        d = b.piece(i).dump()
        assert np.array_equal(d, a.piece(i).dump()), i
        assert np.array_equal(d, orc.dump(4, i)), i
    del b
    if spill_dir:
        assert os.listdir(tmp_path) == [], "spill files left behind"


def test_out_of_core_file_run_writes_the_same_discs(tmp_path):
    g = mx.gen_gaussians(200, 180, 8, 1e-3, seed=2)
    p = tmp_path / "g.raw"
    g.tofile(p)
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    a = mx.GridRun.from_file(str(p), (200, 180), levels=2, threshold=0.01, disc_prefix=str(tmp_path / "a" / "dp_"))
    b = mx.GridRun.from_file(str(p), (200, 180), levels=2, threshold=0.01, disc_prefix=str(tmp_path / "b" / "dp_"),
                             out_of_core=True)
    fa, fb = sorted(os.listdir(tmp_path / "a")), sorted(os.listdir(tmp_path / "b"))
    assert fa == fb and fa
    for f in fa:
        assert (tmp_path / "a" / f).read_bytes() == (tmp_path / "b" / f).read_bytes(), f
    for i in range(a.num_pieces()):
        assert np.array_equal(a.piece(i).dump(), b.piece(i).dump()), i
