"""GPU parity of the file-driven run (mscx_grid_run_file and the mscx_dmsc
command line) against the oracle's GridDataManager order: every piece word
for word after the discs, and the disc files byte for byte as the
reference's collectManifold / write_discs would append them
(grid_datamanager.cpp:397-420, grid_mscomplex.cpp:542-590).  The slab size
of the reader is varied so that grids span one slab, many slabs and a
ragged last slab."""
import os
import subprocess

import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from mscomplex_amd import build as mbuild
from tests.helpers import rng_field

pytestmark = pytest.mark.gpu

BCANC = 4  # the dump's flag of a cancellable (paired) critical point


def _expected_discs(orc, prefix):
    """{file name: bytes} of write_discs over the leaves in piece order:
    for every critical point that is not boundary-cancellable, its ascending
    and descending discs appended to <prefix>asc_(x,y) / des_(x,y)."""
    files = {}
    for i in range(orc.num_leaves()):
        _, cps = oracle.parse_dump(orc.dump(4, i))
        for cp in cps:
            if cp["flags"] & BCANC:
                continue
            x = np.int16(cp["cell"] & 0xFFFF)
            y = np.int16(cp["cell"] >> 16)
            for name, key in (("asc_", "asc_disc"), ("des_", "des_disc")):
                ids = cp[key]
                if not ids:
                    continue
                f = f"{prefix}{name}({x},{y})"
                files[f] = files.get(f, b"") + np.asarray(ids, np.uint32).tobytes()
    return files


def _read_dir(d):
    return {n: open(os.path.join(d, n), "rb").read() for n in sorted(os.listdir(d))}


def _check_pieces(run, orc, label):
    assert run.num_pieces() == orc.num_pieces()
    for i in range(run.num_pieces()):
        a = orc.dump(4, i)
        b = run.piece(i).dump()
        assert a.size == b.size and np.array_equal(a, b), f"{label}: piece {i} differs"


CASES = [
    # (label, field, tiling, threshold, num_parallel, slab bytes or None)
    ("vnoise300x257-l2-7rows", ("vnoise", 300, 257), dict(levels=2), 0.01, 1, 7 * 300 * 4),
    ("white150x130-2x2-1row", ("white", 150, 130), dict(lx=1, ly=1), 0.0, 3, 150 * 4),
    ("gauss512-l0-default", ("gauss", 512, 512), dict(levels=0), 0.0, 0, None),
    ("vnoise1024-4x4-ragged", ("vnoise", 1024, 1024), dict(lx=2, ly=2), 0.01, 0, 100000),
]


def _field(kind, sx, sy):
    if kind == "vnoise":
        return mx.gen_value_noise(sx, sy, period=max(sx, sy) / 4, seed=sx + sy)
    if kind == "gauss":
        return mx.gen_gaussians(sx, sy, 32, 1e-3, seed=1)
    return rng_field(kind, sx, sy, 5)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_file_run_matches_oracle(case, tmp_path, monkeypatch):
    label, (kind, sx, sy), tiling, th, npar, slab = case
    print(f"[file-run] {label}", flush=True)
    data = _field(kind, sx, sy)
    path = tmp_path / "grid.raw"
    data.astype("<f4").tofile(path)
    if slab:
        monkeypatch.setenv("MSCX_READ_SLAB_BYTES", str(slab))
    else:
        monkeypatch.delenv("MSCX_READ_SLAB_BYTES", raising=False)
    out = tmp_path / "discs"
    out.mkdir()
    run = mx.GridRun.from_file(str(path), (sx, sy), threshold=th, num_parallel=npar,
                               disc_prefix=str(out) + "/dp__disc_", **tiling)
    orc = oracle.Pipeline(data, keep=True, threshold=th, **tiling)
    _check_pieces(run, orc, label)
    exp = _expected_discs(orc, "dp__disc_")
    got = _read_dir(out)
    assert sorted(got) == sorted(exp), label
    bad = [n for n in exp if got[n] != exp[n]]
    assert not bad, f"{label}: {len(bad)} disc files differ, e.g. {bad[:3]}"


def test_trailing_bytes_are_ignored(tmp_path):
    data = _field("vnoise", 96, 80)
    path = tmp_path / "grid.raw"
    with open(path, "wb") as f:
        f.write(data.astype("<f4").tobytes())
        f.write(np.full(1000, np.nan, np.float32).tobytes())  # never read
    run = mx.GridRun.from_file(str(path), (96, 80), levels=1, threshold=0.02, discs=False)
    ref = mx.GridRun(data, levels=1, threshold=0.02, discs=False)
    for i in range(ref.num_pieces()):
        assert np.array_equal(run.piece(i).dump(), ref.piece(i).dump())


def test_nan_in_file_is_einval(tmp_path, monkeypatch):
    data = _field("vnoise", 64, 64)
    data[40, 7] = np.nan  # in the third slab of 16 rows
    path = tmp_path / "grid.raw"
    data.tofile(path)
    monkeypatch.setenv("MSCX_READ_SLAB_BYTES", str(16 * 64 * 4))
    with pytest.raises(mx.MscxError) as e:
        mx.GridRun.from_file(str(path), (64, 64))
    assert e.value.code == -1 and "NaN" in str(e.value)
    # the library is usable afterwards
    monkeypatch.delenv("MSCX_READ_SLAB_BYTES")
    data[40, 7] = 0.5
    data.tofile(path)
    mx.GridRun.from_file(str(path), (64, 64), discs=False)


def test_dmsc_command_line(tmp_path):
    """mscx_dmsc -gf <file> -d <x> <y> -l 2 --simp-tresh 0.02 --num-parallel 4
    writes dp__disc_* in its working directory; -st gives the same files."""
    sx, sy = 200, 180
    data = _field("vnoise", sx, sy)
    path = tmp_path / "grid.raw"
    data.tofile(path)
    orc = oracle.Pipeline(data, keep=True, levels=2, threshold=0.02)
    exp = _expected_discs(orc, "dp__disc_")
    for sub, extra in (("mt", ["--num-parallel", "4"]), ("st", ["-st"])):
        d = tmp_path / sub
        d.mkdir()
        print(f"[file-run] mscx_dmsc {sub}", flush=True)
        r = subprocess.run([mbuild.CLI, "-gf", str(path), "-d", str(sx), str(sy), "-l", "2", "--simp-tresh", "0.02",
                            *extra], capture_output=True, text=True, cwd=d, timeout=120)
        assert r.returncode == 0, r.stderr
        assert "pieces 7" in r.stdout
        got = _read_dir(d)
        assert sorted(got) == sorted(exp)
        assert all(got[n] == exp[n] for n in exp), sub


def test_archives_of_device_run(tmp_path):
    """Every piece of a device-computed run survives an archive round trip
    (mscx_complex_save / mscx_complex_load) word for word."""
    data = _field("vnoise", 256, 200)
    run = mx.GridRun(data, lx=1, ly=1, threshold=0.01)
    for i in range(run.num_pieces()):
        p = tmp_path / f"dp_msgraph_{i}"
        run.piece(i).save(p)
        assert np.array_equal(mx.GridMSComplex.load(p).dump(), run.piece(i).dump()), f"piece {i}"
