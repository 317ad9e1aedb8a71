"""Host archives of complexes (mscx_complex_save / mscx_complex_load), the
counterpart of the reference's out-of-core archives
(write_msgraph_to_archive / read_msgraph_from_archive,
grid_datamanager.cpp:91-119): a loaded complex dumps word for word as the
saved one, complexes spilled between the stages of the GridDataManager order
give the oracle's result, and damaged files fail with MSCX_EIO.  Built from
the oracle's per-tile inputs, so no GPU is needed."""
import os

import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from tests.helpers import rng_field, split_pieces
from tests.test_complex_cpu import tile_inputs

EIO = -6


def _leaves(data, **kw):
    sy, sx = data.shape
    out = []
    for rect, ext in split_pieces(sx, sy, **kw):
        crit, fn, conn, pr, codes, own = tile_inputs(data, rect, ext)
        out.append((rect, ext, mx.GridMSComplex.from_critical(rect, ext, crit, fn, conn, pr), codes, own))
    return out


def _spill(c, path):
    c.save(path)
    return mx.GridMSComplex.load(path)


@pytest.mark.parametrize("kw", [dict(levels=2, threshold=0.02), dict(lx=1, ly=1, threshold=0.0)])
def test_spilled_pipeline_matches_oracle(tmp_path, kw):
    """Every piece goes through an archive after each stage that produced
    it (compute, merge up, simplify, merge down) and is used from the loaded
    copy; the final pieces equal the oracle's word for word."""
    data = rng_field("smooth", 120, 100, 3)
    th = kw.pop("threshold")
    leaves = _leaves(data, **kw)
    n = len(leaves)
    pieces = [_spill(c, tmp_path / f"leaf{i}") for i, (_, _, c, _, _) in enumerate(leaves)]
    for j in range(0, 2 * n - 2, 2):
        p = mx.GridMSComplex.merge_up(pieces[j], pieces[j + 1])
        pieces.append(_spill(p, tmp_path / f"up{len(pieces)}"))
    if th > 0:
        pieces[-1].simplify_un_simplify(th)
        pieces[-1] = _spill(pieces[-1], tmp_path / "root")
    for p in range(2 * n - 2, n - 1, -1):
        c = 2 * (p - n)
        pieces[p].merge_down(pieces[c], pieces[c + 1])
        pieces[c] = _spill(pieces[c], tmp_path / f"down{c}")
        pieces[c + 1] = _spill(pieces[c + 1], tmp_path / f"down{c + 1}")
    for i, (rect, ext, _, codes, own) in enumerate(leaves):
        pieces[i].fill_discs_host(rect, ext, codes, own)
        pieces[i] = _spill(pieces[i], tmp_path / f"discs{i}")
    orc = oracle.Pipeline(data, keep=True, threshold=th, **kw)
    assert orc.num_pieces() == len(pieces)
    for i, c in enumerate(pieces):
        assert np.array_equal(orc.dump(4, i), c.dump()), f"piece {i}"


def test_round_trip_keeps_discs(tmp_path):
    data = rng_field("white", 60, 50, 2)
    leaves = _leaves(data, levels=0)
    rect, ext, c, codes, own = leaves[0]
    c.fill_discs_host(rect, ext, codes, own)
    d = c.dump()
    assert d.size > 9 + 8 * c.num_cps()  # discs are in the dump
    back = _spill(c, tmp_path / "c")
    assert np.array_equal(back.dump(), d)
    # and the loaded complex writes the same disc files
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    c.write_discs(str(tmp_path / "a") + "/dp_")
    back.write_discs(str(tmp_path / "b") + "/dp_")
    fa, fb = sorted(os.listdir(tmp_path / "a")), sorted(os.listdir(tmp_path / "b"))
    assert fa == fb and fa
    assert all(open(tmp_path / "a" / f, "rb").read() == open(tmp_path / "b" / f, "rb").read() for f in fa)


def _archive(tmp_path):
    data = rng_field("smooth", 40, 30, 1)
    c = _leaves(data, levels=0)[0][2]
    p = tmp_path / "c"
    c.save(p)
    return p, bytearray(open(p, "rb").read())


def _load_fails(p):
    with pytest.raises(mx.MscxError) as e:
        mx.GridMSComplex.load(p)
    assert e.value.code == EIO
    return str(e.value)


def test_missing_archive(tmp_path):
    assert "cannot read" in _load_fails(tmp_path / "none")


def test_foreign_file(tmp_path):
    p, b = _archive(tmp_path)
    b[0:8] = b"NOTMSCX!"
    open(p, "wb").write(b)
    assert "not a complex archive" in _load_fails(p)


def test_other_version(tmp_path):
    p, b = _archive(tmp_path)
    b[8] = 2
    open(p, "wb").write(b)
    assert "version" in _load_fails(p)


@pytest.mark.parametrize("cut", [1, 4, 100])
def test_truncated_archive(tmp_path, cut):
    p, b = _archive(tmp_path)
    open(p, "wb").write(b[:-cut])
    assert "truncated" in _load_fails(p)


@pytest.mark.parametrize("where", [24, 40, -9])
def test_flipped_bit(tmp_path, where):
    p, b = _archive(tmp_path)
    b[where] ^= 0x10
    open(p, "wb").write(b)
    assert "corrupt" in _load_fails(p)
