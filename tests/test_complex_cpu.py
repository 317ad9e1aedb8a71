"""CPU parity of the product's complex algebra (libmscx.so host code:
writeout_connectivity, merge_up, simplify/un_simplify, merge_down,
postMergeFillDiscs) against the oracle pipeline.

The per-tile inputs (critical lists, owners, codes) come from the oracle's
GridDataset::work(), so these tests need no GPU; the GPU-produced inputs are
checked separately (test_gpu_tile.py, test_gpu_pipeline.py).  Every piece's
complex is compared word for word after the whole GridDataManager order
(grid_datamanager.cpp:643-660)."""
import numpy as np
import pytest

import mscomplex_amd as mx
import oracle
from tests.helpers import rng_field, split_pieces, window

DIRS = {(1, 0): 1, (-1, 0): 2, (0, 1): 3, (0, -1): 4}


def tile_inputs(data, rect, ext):
    """Oracle work() on one tile -> (crit, fn, conn4, pair, codes, own)."""
    win = window(data, ext)
    flags, pairs, own, crit = oracle.tile_work(win, rect, ext)
    _, cps = oracle.parse_dump(oracle.tile_complex(win, rect, ext))
    K = crit.size
    idx = {int(c): k for k, c in enumerate(crit)}
    fn = np.array([cp["fn"] for cp in cps], np.uint32).view(np.float32)
    pr = np.array([cp["pair"] for cp in cps], np.uint32)
    ch, cw = own.shape
    x0, y0 = ext[0], ext[1]

    def own_idx(x, y):
        if x < ext[0] or y < ext[1] or x > ext[2] or y > ext[3]:
            return 0xFFFFFFFF
        o = int(own[y - y0, x - x0])
        return 0xFFFFFFFF if o == 0xFFFFFFFF else idx[o]

    conn = np.full((K, 4), 0xFFFFFFFF, np.uint32)
    for k, c in enumerate(crit):
        x, y = mx.unpack(c)
        dx, dy = x & 1, y & 1
        if dx == dy:
            continue
        conn[k] = (own_idx(x + dx, y + dy), own_idx(x - dx, y - dy), own_idx(x + dy, y + dx),
                   own_idx(x - dy, y - dx))
    codes = (flags.astype(np.uint8) & 3)
    ys, xs = np.nonzero(pairs != 0xFFFFFFFF)
    for y, x in zip(ys, xs):
        px, py = mx.unpack(pairs[y, x])
        codes[y, x] |= DIRS[(px - (x + x0), py - (y + y0))] << 2
    return crit, fn, conn, pr, codes, own


def product_pipeline(data, levels=0, lx=0, ly=0, threshold=0.0):
    sy, sx = data.shape
    leaves = split_pieces(sx, sy, levels=levels, lx=lx, ly=ly)
    n = len(leaves)
    pieces, inputs = [], []
    for rect, ext in leaves:
        crit, fn, conn, pr, codes, own = tile_inputs(data, rect, ext)
        inputs.append((codes, own))
        pieces.append(mx.GridMSComplex.from_critical(rect, ext, crit, fn, conn, pr))
    for j in range(0, 2 * n - 2, 2):
        pieces.append(mx.GridMSComplex.merge_up(pieces[j], pieces[j + 1]))
    if threshold > 0:
        pieces[-1].simplify_un_simplify(threshold)
    # the reference's single-threaded merge-down order: the internal levels
    # from the highest piece down (mergePiecesDown), then the leaves'
    # parents upwards (finalMergeDownPiecesInRange)
    for p in list(range(2 * n - 2, n + n // 2 - 1, -1)) + list(range(n, n + n // 2)):
        c = 2 * (p - n)
        pieces[p].merge_down(pieces[c], pieces[c + 1])
    for i, (rect, ext) in enumerate(leaves):
        codes, own = inputs[i]
        pieces[i].fill_discs_host(rect, ext, codes, own)
    return pieces


def compare(data, **kw):
    orc = oracle.Pipeline(data, keep=True, **kw)
    mine = product_pipeline(data, **kw)
    assert orc.num_pieces() == len(mine)
    for i, m in enumerate(mine):
        a = orc.dump(4, i)
        b = m.dump()
        assert a.size == b.size and np.array_equal(a, b), f"piece {i} differs ({a.size} vs {b.size} words)"


TILINGS = [dict(levels=0), dict(levels=1), dict(levels=2), dict(levels=3), dict(lx=1, ly=1), dict(lx=2, ly=2)]


@pytest.mark.parametrize("kind", ["white", "smooth"])
@pytest.mark.parametrize("tiling", TILINGS)
@pytest.mark.parametrize("threshold", [0.0, 0.02, 0.3])
def test_pipeline_matches_oracle(kind, tiling, threshold):
    data = rng_field(kind, 37, 29, seed=11)
    compare(data, threshold=threshold, **tiling)


# A 3-level field at a threshold above 0.1 makes the reference's
# cancelPairs multiply strangulated arcs without bound (the oracle does not
# finish); ties are checked at thresholds where it terminates.
@pytest.mark.parametrize("tiling", TILINGS)
@pytest.mark.parametrize("threshold", [0.0, 0.02, 0.1])
def test_pipeline_matches_oracle_ties(tiling, threshold):
    data = rng_field("ties", 37, 29, seed=11)
    compare(data, threshold=threshold, **tiling)


@pytest.mark.parametrize("seed", range(6))
def test_gaussian_fields(seed):
    data = mx.gen_gaussians(48, 40, 8, 1e-3, seed=seed)
    compare(data, lx=1, ly=1, threshold=0.05)


def test_value_noise_4x4():
    data = mx.gen_value_noise(96, 96, period=24, seed=4)
    compare(data, lx=2, ly=2, threshold=0.01)


def test_quantised_white_ties_heavy():
    data = mx.gen_white(64, 64, levels=8, seed=2)
    compare(data, lx=1, ly=2, threshold=0.1)


def test_dense_cancellation_fails_cleanly():
    data = rng_field("white", 150, 130, 5)
    rect = (0, 0, 2 * 149, 2 * 129)
    crit, fn, conn, pr, codes, own = tile_inputs(data, rect, rect)
    c = mx.GridMSComplex.from_critical(rect, rect, crit, fn, conn, pr)
    with pytest.raises(mx.MscxError) as e:
        c.simplify_un_simplify(0.1)
    assert e.value.code == -4


# Tilings finer than the grid: some pieces degenerate and the reference's
# own invariant checks (logic_error) fire.  The product must fail where the
# reference's single-threaded order fails first, with the same message, and
# succeed (word for word) where it succeeds.
DEGENERATE = [(5, 5, dict(levels=3)), (2, 2, dict(levels=1)), (2, 9, dict(levels=2)), (9, 2, dict(levels=1)),
              (3, 3, dict(levels=4)), (2, 2, dict(levels=0)), (3, 3, dict(lx=1, ly=1)), (2, 5, dict(lx=1, ly=1)),
              (6, 6, dict(lx=2, ly=2)), (4, 17, dict(levels=3)), (3, 40, dict(levels=4)), (40, 3, dict(lx=3, ly=0))]


def oracle_outcome(data, **kw):
    try:
        return oracle.Pipeline(data, keep=True, **kw), None
    except RuntimeError as e:
        return None, str(e)


TOO_FINE = "tiling is finer than the grid"


@pytest.mark.parametrize("case", DEGENERATE, ids=[f"{a}x{b}-{kw}" for a, b, kw in DEGENERATE])
def test_degenerate_tilings(case):
    sx, sy, kw = case
    data = rng_field("white", sx, sy, 1)
    orc, msg = oracle_outcome(data, threshold=0.0, **kw)
    if msg is not None and msg.startswith(TOO_FINE):
        # a piece's ext rect would leave the grid: refused before any data
        # is read (mscx_grid_run checks the tiling before the device)
        with pytest.raises(mx.MscxError) as e:
            mx.GridRun(data, threshold=0.0, **kw)
        assert e.value.code == -1 and TOO_FINE in str(e.value)
    elif msg is not None:
        with pytest.raises(mx.MscxError) as e:
            product_pipeline(data, threshold=0.0, **kw)
        assert e.value.code == -3 and str(e.value).endswith(msg), (str(e.value), msg)
    else:
        mine = product_pipeline(data, threshold=0.0, **kw)
        assert len(mine) == orc.num_pieces()
        for i, c in enumerate(mine):
            assert np.array_equal(orc.dump(4, i), c.dump()), f"piece {i}"


def test_large_merges_take_the_threaded_path():
    # a root of more than 16384 nodes: merge_up looks cells up and fills
    # the union on several threads and builds its cell map on one thread
    # while the others build the arc sets by ranges of destination nodes;
    # the result must not depend on it
    data = rng_field("white", 300, 260, 4)
    compare(data, lx=1, ly=1, threshold=0.0)
    compare(data, levels=2, threshold=0.01)
