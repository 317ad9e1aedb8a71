"""The C ABI library loads without a GPU and exports every function that
include/mscx.h declares; host-only entry points behave as documented."""
import os
import re

import numpy as np
import pytest

import mscomplex_amd as mx
from mscomplex_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "mscx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mscx_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert not [n for n in names if n not in _lib._SIGS]


def test_errors_are_status_codes():
    with pytest.raises(mx.MscxError) as e:
        mx.GridMSComplex.from_critical((0, 0, 4, 4), (0, 0, 4, 4), np.array([0], np.uint32),
                                       np.zeros(1, np.float32), np.array([[5, 0, 0, 0]], np.uint32),
                                       np.array([0xFFFFFFFF], np.uint32))
    assert e.value.code == -1 and "out of range" in str(e.value)


def test_generators_are_deterministic():
    a = mx.gen_value_noise(64, 48, seed=3)
    b = mx.gen_value_noise(64, 48, seed=3)
    assert a.dtype == np.float32 and a.shape == (48, 64) and np.array_equal(a, b)
    g = mx.gen_gaussians(50, 40, 8, 1e-3, seed=1)
    assert np.isfinite(g).all() and g.std() > 0
    w = mx.gen_white(30, 20, levels=256, seed=2)
    assert np.array_equal(np.floor(w * 256) / 256, w)


def test_write_discs_format(tmp_path):
    # a complex built from host data; discs appended in the reference's raw format
    c = mx.GridMSComplex.from_critical((0, 0, 2, 2), (0, 0, 2, 2), np.array([0], np.uint32),
                                       np.zeros(1, np.float32), np.full((1, 4), 0xFFFFFFFF, np.uint32),
                                       np.array([0xFFFFFFFF], np.uint32))
    assert c.num_cps() == 1
    c.write_discs(str(tmp_path) + "/dp_")
    assert os.listdir(tmp_path) == []  # empty discs write nothing


def test_autotune_prefers_the_coarser_tiling_within_the_margin():
    pick = mx._pick_tiling
    assert pick([], 0.05) == (0, 0)
    assert pick([(0.5709, (0, 0)), (0.5702, (1, 1)), (0.6376, (2, 2))], 0.05) == (0, 0)
    assert pick([(1.0, (0, 0)), (0.90, (1, 1)), (0.80, (2, 2))], 0.05) == (2, 2)
    assert pick([(1.0, (0, 0)), (0.90, (1, 1)), (0.88, (2, 2))], 0.05) == (1, 1)


def test_block_rows_table_is_validated():
    # no device work: the table only selects the kernels' block height
    mx.set_block_rows([(127, 16), (511, 32)])
    mx.set_block_rows(None)
    for bad in ([(100, 24)], [(511, 16), (127, 32)], [(i, 16) for i in range(1, 18)]):
        with pytest.raises(mx.MscxError):
            mx.set_block_rows(bad)
    mx.set_block_rows(None)
