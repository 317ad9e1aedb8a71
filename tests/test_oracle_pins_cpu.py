"""Pins of the oracle's merge tree, cancellation, simplification and
manifold functions (merge_up, cancelPairs, uncancel_pairs, simplify,
un_simplify, merge_down, postMergeFillDiscs) by facts the mathematics fixes,
not by re-typing the reference's steps:

  (i)   tiling independence: the root after merge_up at threshold 0,
        canonicalised by cell id, is the complex of the untiled grid;
  (ii)  Morse homology of a disc over Z/2 on the surviving complex, before
        and after simplify_un_simplify: d1 d2 = 0, rank d1 = #min - 1,
        rank d2 = #max, every saddle keeps two descending and two (one on
        the boundary) ascending paths;
  (iii) the surviving extrema's discs partition each leaf rect's vertices
        and quads, and (untiled, threshold 0) each minimum's disc is exactly
        its gradient basin found by brute force; with cancellations every
        basin lies inside one disc;
  (iv)  after merge_down, each leaf's surviving nodes are the root's
        survivors inside its rect;
  (v)   simplify's range starts from FLT_MIN (grid_mscomplex.cpp:399): an
        all-negative integer field cancels the same pairs as its exact
        positive shift at thresholds rescaled by the ranges that rule gives.

Each check is shown to fail on a plausible mistake planted in the oracle
(oracle.plant)."""
import numpy as np
import pytest

import oracle
from tests import pins
from tests.helpers import rng_field

FIELDS = ["white", "smooth", "ties"]
TILINGS = [dict(levels=1), dict(levels=3), dict(lx=1, ly=1), dict(lx=2, ly=1), dict(lx=1, ly=2)]


def live(dump):
    return (pins.nodes(dump)["flags"] & pins.F_CANCELLED) == 0


def root(pl):
    return pl.num_pieces() - 1


def simplify_survivors(pl):
    """Root nodes alive after merge_up that simplify did not cancel (the
    boundary-cancellable flag it sets marks its pairs)."""
    r = root(pl)
    f1 = pins.nodes(pl.dump(1, r))["flags"]
    f2 = pins.nodes(pl.dump(2, r))["flags"]
    newly = ((f2 & pins.F_BCANC) != 0) & ((f1 & pins.F_BCANC) == 0)
    return ((f1 & pins.F_CANCELLED) == 0) & ~newly


# ------------------------------------------------------------------ (i)

@pytest.mark.parametrize("kind", FIELDS + ["signedzero"])
@pytest.mark.parametrize("shape", [(33, 27), (17, 45)])
def test_merge_up_root_equals_untiled_complex(kind, shape):
    sx, sy = shape
    data = rng_field(kind, sx, sy, seed=sx * 3 + sy)
    flat = oracle.Pipeline(data, levels=0)
    c0 = pins.canon(flat.dump(1, 0))
    for til in TILINGS:
        pl = oracle.Pipeline(data, threshold=0.0, **til)
        assert np.array_equal(pins.canon(pl.dump(1, root(pl))), c0), til


def test_merge_up_pin_catches_missing_dedup():
    data = rng_field("smooth", 33, 27, seed=4)
    c0 = pins.canon(oracle.Pipeline(data, levels=0).dump(1, 0))
    with oracle.plant(1, stop_after=1):
        pl = oracle.Pipeline(data, levels=2)
    assert not np.array_equal(pins.canon(pl.dump(1, root(pl))), c0)


# ------------------------------------------------------------------ (ii)

@pytest.mark.parametrize("kind", FIELDS)
@pytest.mark.parametrize("threshold", [0.0, 0.02, 0.05])
@pytest.mark.parametrize("tiling", [dict(levels=0), dict(levels=2), dict(lx=1, ly=1)])
def test_disc_homology_before_and_after_simplify(kind, threshold, tiling):
    data = rng_field(kind, 40, 30, seed=11)
    pl = oracle.Pipeline(data, threshold=threshold, **tiling)
    r = root(pl)
    before = pins.assert_disc_homology(pl.dump(1, r), live(pl.dump(1, r)), "merged")
    after = pins.assert_disc_homology(pl.dump(2, r), simplify_survivors(pl), "simplified")
    if threshold == 0.0:
        assert before == after
    else:
        assert after["saddle"] < before["saddle"] or kind == "ties"


def _homology_fails(dump, mask):
    try:
        pins.assert_disc_homology(dump, mask)
    except AssertionError:
        return True
    return False


def test_homology_pin_catches_dropped_arc_multiplicity():
    data = rng_field("white", 40, 30, seed=11)
    with oracle.plant(2, stop_after=2):
        pl = oracle.Pipeline(data, levels=0, threshold=0.05)
    assert _homology_fails(pl.dump(2, 0), simplify_survivors(pl))


# ------------------------------------------------------------------ (iii)

def _basins(data):
    """Gradient basin (critical minimum) of every vertex and the maximum
    above every quad of the untiled grid, by chasing the oracle's pairs."""
    sy, sx = data.shape
    f, p, o, c = oracle.tile_work(data, (0, 0, 2 * (sx - 1), 2 * (sy - 1)))
    px = (p & 0xFFFF).astype(np.int64)
    py = (p >> 16).astype(np.int64)
    ch, cw = f.shape
    # next cell of a vertex / quad across its paired edge (itself if critical)
    yy, xx = np.mgrid[0:ch, 0:cw]
    stay = ((f & 2) != 0) | (((xx + yy) & 1) == 1)  # critical cells and edges
    nx = np.where(stay, xx, 2 * px - xx)
    ny = np.where(stay, yy, 2 * py - yy)
    nxt = ny * cw + nx
    for _ in range(int(np.ceil(np.log2(ch * cw))) + 2):
        nxt = nxt.reshape(-1)[nxt]
    return nxt, cw


@pytest.mark.parametrize("kind", FIELDS)
@pytest.mark.parametrize("threshold", [0.0, 0.03])
@pytest.mark.parametrize("tiling", [dict(levels=0), dict(levels=2), dict(lx=1, ly=1), dict(lx=2, ly=2)])
def test_discs_partition_each_leaf(kind, threshold, tiling):
    data = rng_field(kind, 41, 33, seed=5)
    pl = oracle.Pipeline(data, threshold=threshold, **tiling)
    nxt, cw = _basins(data)
    for leaf in range(pl.num_leaves()):
        rect, _ = pl.rects(leaf)
        d = pl.dump(4, leaf)
        pins.assert_disc_partition(d, rect, (leaf, rect))
        # every gradient basin lies inside one surviving extremum's disc
        _, cps = oracle.parse_dump(d)
        for key, dim in (("asc_disc", 0), ("des_disc", 2)):
            ext_idx = [i for i, cp in enumerate(cps) if pins.dims([cp["cell"]])[0] == dim]
            owner = {}
            for i in ext_idx:
                for v in cps[i][key]:
                    owner[v] = i
            by_basin = {}
            for v, i in owner.items():
                x, y = v & 0xFFFF, v >> 16
                by_basin.setdefault(int(nxt[y, x]), set()).add(i)
            assert all(len(s) == 1 for s in by_basin.values()), (leaf, key)
            if tiling == dict(levels=0) and threshold == 0.0:
                # untiled, nothing cancelled: a disc is exactly one basin
                for i in ext_idx:
                    cp = cps[i]
                    if cp[key]:
                        x, y = cp["cell"] & 0xFFFF, cp["cell"] >> 16
                        exp = {int(v) for v in np.flatnonzero(nxt.reshape(-1) == y * cw + x)
                               if ((v % cw) & 1) == dim // 2 and ((v // cw) & 1) == dim // 2}
                        got = {(v >> 16) * cw + (v & 0xFFFF) for v in cp[key]}
                        assert got == exp, (i, key)


def test_disc_pin_catches_bad_uncancel_rebuild():
    data = rng_field("smooth", 41, 33, seed=5)
    caught = False
    with oracle.plant(3):
        try:
            pl = oracle.Pipeline(data, lx=1, ly=1, threshold=0.03)
        except RuntimeError:
            caught = True  # the reference's own invariant check fires
        else:
            for leaf in range(pl.num_leaves()):
                rect, _ = pl.rects(leaf)
                caught |= any(pins.discs(pl.dump(4, leaf), rect).values())
    assert caught


# ------------------------------------------------------------------ (iv)

@pytest.mark.parametrize("kind", FIELDS)
@pytest.mark.parametrize("threshold", [0.0, 0.02, 0.05])
@pytest.mark.parametrize("tiling", [dict(levels=2), dict(levels=3), dict(lx=1, ly=1), dict(lx=2, ly=1)])
def test_merge_down_leaf_nodes_are_root_survivors(kind, threshold, tiling):
    data = rng_field(kind, 40, 30, seed=21)
    pl = oracle.Pipeline(data, threshold=threshold, **tiling)
    rn = pins.nodes(pl.dump(2, root(pl)))
    rs = rn["cell"][(rn["flags"] & pins.F_BCANC) == 0]
    for leaf in range(pl.num_leaves()):
        rect, _ = pl.rects(leaf)
        n = pins.nodes(pl.dump(3, leaf))
        ls = n["cell"][((n["flags"] & pins.F_BCANC) == 0) & pins.cells_in(n["cell"], rect)]
        assert np.array_equal(np.sort(ls), np.sort(rs[pins.cells_in(rs, rect)])), leaf


# ------------------------------------------------------------------ (v)

def _int_field():
    r = np.random.default_rng(5)
    sy, sx = 30, 40
    y, x = np.mgrid[0:sy, 0:sx]
    base = np.round(6 * np.sin(x / 4.0) * np.cos(y / 5.0) + 2 * r.random((sy, sx))).astype(np.float32)
    neg = base - base.max() - 3  # every value <= -3
    return neg, neg + 30  # exact shift: integers, persistences unchanged


def _fn(dump):
    return pins.nodes(dump)["fn"].view(np.float32)


@pytest.mark.parametrize("absolute", [0.6, 1.6])
def test_flt_min_range_rule(absolute):
    neg, pos = _int_field()
    vn = _fn(oracle.Pipeline(neg, levels=0).dump(1, 0))
    vp = _fn(oracle.Pipeline(pos, levels=0).dump(1, 0))
    assert vn.max() < 0 < vp.min()
    # max_val starts at FLT_MIN, so an all-negative field's range is -min
    range_neg = float(np.float32(np.finfo(np.float32).tiny) - vn.min())
    range_pos = float(vp.max() - vp.min())
    a = oracle.Pipeline(neg, levels=0, threshold=absolute / range_neg)
    b = oracle.Pipeline(pos, levels=0, threshold=absolute / range_pos)
    for st in (2, 3, 4):
        assert np.array_equal(pins.without_fn(a.dump(st, 0)), pins.without_fn(b.dump(st, 0))), st
    nb = pins.nodes(b.dump(2, 0))
    assert ((nb["flags"] & pins.F_BCANC) != 0).sum() > 0
    # the mistake of starting max_val at the lowest float changes the result
    with oracle.plant(4, stop_after=2):
        c = oracle.Pipeline(neg, levels=0, threshold=absolute / range_neg)
    assert not np.array_equal(pins.without_fn(c.dump(2, 0)), pins.without_fn(a.dump(2, 0)))
