"""The engine's heap (csrc/heap.hpp) must pop equal-key elements in exactly
the order of std::priority_queue, because the reference's simplification
(grid_mscomplex.cpp:387-510) resolves equal persistences by that order."""
import ctypes
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM = os.path.join(HERE, "_shim", "libheap.so")


@pytest.fixture(scope="module")
def shim():
    os.makedirs(os.path.dirname(SHIM), exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", SHIM,
                    os.path.join(HERE, "heap_shim.cpp")], check=True)
    L = ctypes.CDLL(SHIM)
    L.heap_compare.restype = ctypes.c_long
    L.heap_compare.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return L


@pytest.mark.parametrize("nkeys", [1, 2, 3, 17, 1000, 1 << 30])
@pytest.mark.parametrize("push_pct", [30, 50, 70])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pop_order_matches_priority_queue(shim, nkeys, push_pct, seed):
    assert shim.heap_compare(seed * 7919 + nkeys, 200000, nkeys, push_pct) == 0
