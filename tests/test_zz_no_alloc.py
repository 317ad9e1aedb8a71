"""mscx_tile_work_noalloc (the entry point for callers that cannot allocate
device memory) against mscx_tile_work on the same tiles, through the C ABI:
the critical list, cell values, saddle incidence and pair indices word for
word, on two calls in a row (the second on a smaller tile, so the shared
scratch holds the first call's leftovers), plus its two refusals (a tile
beyond the scratch, a too-small output capacity)."""
import ctypes

import numpy as np
import pytest

import mscomplex_amd as mx
from mscomplex_amd._lib import lib
from tests.helpers import split_pieces, window

pytestmark = pytest.mark.gpu


def _reference(data, rect, ext):
    t = mx.GridDataset(rect, ext)
    t.init(window(data, ext))
    t.work()
    crit = t.arrays(flags=False, pairs=False, own=False)[3]
    fn, conn, pair = t.critical_data()
    return crit, fn, conn.reshape(-1), pair


def _noalloc(dev, side, rect, ext, kcap):
    outs = [mx.DeviceBuffer(max(4 * kcap, 4)) for _ in range(2)] + [mx.DeviceBuffer(max(16 * kcap, 16)),
                                                                     mx.DeviceBuffer(max(4 * kcap, 4))]
    K = ctypes.c_uint32(0)
    r = (ctypes.c_int32 * 4)(*rect)
    e = (ctypes.c_int32 * 4)(*ext)
    origin = dev.ptr + 4 * ((ext[1] // 2) * side + ext[0] // 2)
    rc = lib().mscx_tile_work_noalloc(r, e, ctypes.c_void_p(origin), side, kcap, *[ctypes.c_void_p(o.ptr) for o in outs],
                                      ctypes.byref(K))
    if rc != 0:
        return rc, K.value, None
    k = K.value
    crit = outs[0].download(np.empty(k, np.uint32))
    fn = outs[1].download(np.empty(k, np.float32))
    conn = outs[2].download(np.empty(4 * k, np.uint32))
    pair = outs[3].download(np.empty(k, np.uint32))
    return rc, k, (crit, fn, conn, pair)


def _same(a, b):
    for x, y in zip(a, b):
        assert x.dtype == y.dtype and x.shape == y.shape
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.timeout(300)
def test_noalloc_matches_tile_work_twice():
    side = 2048
    data = mx.gen_value_noise(side, side, seed=5, period=512)
    dev = mx.DeviceBuffer.from_array(data)
    big = split_pieces(side, side, lx=1, ly=1)[3]     # a 1024^2 leaf with its halo
    small = split_pieces(side, side, lx=3, ly=3)[17]  # a 256^2 leaf, elsewhere in the grid
    for rect, ext in (big, small):
        print(f"[noalloc] rect={rect} ext={ext}", flush=True)
        ref = _reference(data, rect, ext)
        rc, k, got = _noalloc(dev, side, rect, ext, kcap=ref[0].size + 7)
        assert rc == 0, lib().mscx_last_error().decode()
        assert k == ref[0].size
        _same(got, ref)


@pytest.mark.timeout(300)
def test_noalloc_refusals():
    side = 2048
    data = mx.gen_white(side, side, seed=9)
    dev = mx.DeviceBuffer.from_array(data)
    rect = ext = (0, 0, 2 * (side - 1), 2 * 1100)  # 2048 x 1101 vertices: beyond the scratch
    rc, _, _ = _noalloc(dev, side, rect, ext, kcap=16)
    assert rc == -1 and "too large" in lib().mscx_last_error().decode()
    rect, ext = split_pieces(side, side, lx=3, ly=3)[5]
    ref = _reference(data, rect, ext)
    rc, k, _ = _noalloc(dev, side, rect, ext, kcap=ref[0].size - 1)
    assert rc == -1 and k == ref[0].size and "capacity" in lib().mscx_last_error().decode()
    rc, k, got = _noalloc(dev, side, rect, ext, kcap=ref[0].size)
    assert rc == 0
    _same(got, ref)
