"""The header-only C++ drop-in (include/mscx_dataset.hpp) driven in the
reference's GridDataManager order by mscomplex_amd/mscx_dataset_driver: every
piece's complex equals the oracle's word for word, and the host arrays the
header fills (m_cell_flags, m_cell_pairs, m_cell_own, m_critical_cells) are
the reference's, in its [x][y] (y fastest) layout."""
import os
import subprocess

import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from mscomplex_amd import build as mbuild
from tests.helpers import split_pieces, window

pytestmark = pytest.mark.gpu


def _read_pieces(path):
    b = open(path, "rb").read()
    n = int(np.frombuffer(b[:4], np.uint32)[0])
    at, out = 4, []
    for _ in range(n):
        ln = int(np.frombuffer(b[at:at + 8], np.uint64)[0])
        at += 8
        out.append(np.frombuffer(b[at:at + 4 * ln], np.uint32))
        at += 4 * ln
    return out


def _unpack_pairs(words):
    x = (words & 0xFFFF).astype(np.uint16).view(np.int16)
    y = (words >> 16).astype(np.uint16).view(np.int16)
    return np.stack([x, y], -1)


@pytest.mark.parametrize("field,levels,threshold", [("vnoise", 2, 0.01), ("gauss", 0, 0.0), ("white", 3, 0.02)])
def test_driver_matches_oracle(tmp_path, field, levels, threshold):
    sx, sy = 180, 150
    if field == "vnoise":
        g = mx.gen_value_noise(sx, sy, seed=3, period=48)
    elif field == "gauss":
        g = mx.gen_gaussians(sx, sy, 12, 1e-3, seed=3)
    else:
        g = mx.gen_white(sx, sy, seed=3)
    raw = tmp_path / "g.raw"
    g.tofile(raw)
    out, arr = tmp_path / "pieces.bin", tmp_path / "arrays.bin"
    print(f"[hpp] {field} levels={levels} th={threshold}", flush=True)
    r = subprocess.run([mbuild.DRIVER, str(raw), str(sx), str(sy), str(levels), str(threshold), str(out), str(arr)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = _read_pieces(out)
    orc = oracle.Pipeline(g, levels=levels, threshold=threshold)
    assert len(got) == orc.num_pieces()
    for i, d in enumerate(got):
        assert np.array_equal(d, orc.dump(4, i)), f"piece {i}"
    # leaf 0's arrays in the reference's [x][y] layout
    rect, ext = split_pieces(sx, sy, levels=levels)[0]
    f, p, o, c = oracle.tile_work(window(g, ext), rect, ext)  # (ch, cw) images, x fastest
    ch, cw = f.shape
    b = open(arr, "rb").read()
    n = cw * ch
    flags = np.frombuffer(b[:n], np.int8).reshape(cw, ch)
    pairs = np.frombuffer(b[n:5 * n], np.uint32).reshape(cw, ch)
    own = np.frombuffer(b[5 * n:9 * n], np.uint32).reshape(cw, ch)
    crit = np.frombuffer(b[9 * n:], np.uint32)
    assert np.array_equal(flags, f.T)
    assert np.array_equal(pairs, p.T)
    assert np.array_equal(own, o.T)
    assert np.array_equal(crit, c)
    # spot check of the layout: [x][y] with y fastest
    x, y = cw // 3, ch // 2
    assert tuple(_unpack_pairs(pairs[x:x + 1, y])[0]) == tuple(_unpack_pairs(p[y:y + 1, x])[0])
