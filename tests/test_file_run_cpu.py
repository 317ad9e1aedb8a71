"""The file-driven run (mscx_grid_run_file) and the mscx_dmsc command line
without a GPU: argument checks and file errors that come before the device
is touched, and the reference's flag meanings and defaults (main.cpp:106-205,
grid_datamanager.cpp:627-633)."""
import json
import os
import subprocess

import numpy as np
import pytest

import mscomplex_amd as mx
from mscomplex_amd import build as mbuild

CLI = mbuild.CLI


@pytest.fixture(scope="module", autouse=True)
def _built():
    # the command line is built with the library; build both if it is missing
    if not os.path.exists(CLI) or not os.path.exists(mbuild.LIB):
        mbuild.build()


def _cli(*args, cwd=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


def _settings(*args):
    r = _cli(*args, "--print-args")
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_missing_file_is_eio(tmp_path):
    with pytest.raises(mx.MscxError) as e:
        mx.GridRun.from_file(str(tmp_path / "none.raw"), (16, 16))
    assert e.value.code == -6 and "unable to read stream" in str(e.value)


def test_short_file_is_eio(tmp_path):
    p = tmp_path / "g.raw"
    np.zeros((9, 16), np.float32).tofile(p)  # one row short of 16 x 10
    with pytest.raises(mx.MscxError) as e:
        mx.GridRun.from_file(str(p), (16, 10))
    assert e.value.code == -6 and "fewer than" in str(e.value)


def test_directory_is_eio(tmp_path):
    with pytest.raises(mx.MscxError) as e:
        mx.GridRun.from_file(str(tmp_path), (16, 16))
    assert e.value.code == -6


@pytest.mark.parametrize("kw", [dict(num_parallel=-1), dict(threshold=-0.5), dict(levels=13),
                                dict(discs=False, disc_prefix="dp_")])
def test_bad_arguments_are_einval(tmp_path, kw):
    p = tmp_path / "g.raw"
    np.zeros((16, 16), np.float32).tofile(p)
    with pytest.raises(mx.MscxError) as e:
        mx.GridRun.from_file(str(p), (16, 16), **kw)
    assert e.value.code == -1


def test_bad_size_is_einval(tmp_path):
    p = tmp_path / "g.raw"
    np.zeros((16, 16), np.float32).tofile(p)
    with pytest.raises(mx.MscxError) as e:
        mx.GridRun.from_file(str(p), (1, 16))
    assert e.value.code == -1


def test_cli_defaults_follow_the_reference():
    s = _settings("-gf", "a.raw", "-d", "64", "32")
    # parse_grid: levels 1, num_parallel 1 (=> single-threaded), no simplification
    assert s == {"file": "a.raw", "size": [64, 32], "levels": 1, "lx": 0, "ly": 0, "simp_tresh": 0.0,
                 "num_parallel": 1, "single_threaded": True, "write_discs": True, "out_of_core": False}


def test_cli_flags():
    s = _settings("-gf", "dir/a.raw", "-d", "4096", "4096", "-l", "3", "--simp-tresh", "0.01", "--num-parallel", "8",
                  "-ocl", "--out-of-core")
    assert s["levels"] == 3 and s["simp_tresh"] == 0.01 and s["num_parallel"] == 8 and not s["single_threaded"]
    assert s["out_of_core"]
    # -st wins over --num-parallel, as in GridDataManager's constructor
    s = _settings("-st", "--num-parallel", "8", "-gf", "a.raw", "-d", "8", "8")
    assert s["num_parallel"] == 1 and s["single_threaded"]
    s = _settings("-gf", "a.raw", "-d", "8", "8", "--lx", "2", "--ly", "1", "--no-write")
    assert (s["lx"], s["ly"], s["write_discs"]) == (2, 1, False)


@pytest.mark.parametrize("args", [[], ["-gf", "a.raw"], ["-d", "8", "8"], ["-gf", "a.raw", "-d", "8"],
                                  ["-gf", "a.raw", "-d", "8", "8", "--simp-tresh", "1.5"],
                                  ["-gf", "a.raw", "-d", "1", "8"], ["-gf", "a.raw", "-d", "8", "8", "--bogus"],
                                  ["-gf", "a.raw", "-d", "8", "8", "--num-parallel", "0"]])
def test_cli_usage_errors(args):
    r = _cli(*args)
    assert r.returncode == 2 and "Usage:" in r.stderr


def test_cli_missing_file(tmp_path):
    r = _cli("-gf", str(tmp_path / "none.raw"), "-d", "8", "8", cwd=tmp_path)
    assert r.returncode == 1 and "unable to read stream" in r.stderr
    assert os.listdir(tmp_path) == []


def test_tiling_finer_than_the_grid_is_einval(tmp_path):
    # 2 x 2 vertices split into 2 slabs (the reference's default -l 1): the
    # upper piece's ext rect would reach past the grid
    p = tmp_path / "g.raw"
    np.zeros((2, 2), np.float32).tofile(p)
    with pytest.raises(mx.MscxError) as e:
        mx.GridRun.from_file(str(p), (2, 2))
    assert e.value.code == -1 and "finer than the grid" in str(e.value)
    r = _cli("-gf", str(p), "-d", "2", "2", cwd=tmp_path)
    assert r.returncode == 1 and "finer than the grid" in r.stderr


def test_dataset_driver_usage():
    # the GridDataManager-order driver over include/mscx_dataset.hpp is
    # built with the library and refuses bad arguments without a GPU
    r = subprocess.run([mbuild.DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
