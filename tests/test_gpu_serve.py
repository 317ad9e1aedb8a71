"""Served form (mscx_server_*): continuous batching over a request queue
with HBM-budget admission and resident grids for repeat requests.  Every
result must equal the plain single-grid run (mscx_grid_run) word for word,
alone and in a mix, and the oracle where it is cheap; repeat requests served
from a resident grid must equal a fresh run at their own threshold."""
import threading
import time

import numpy as np
import pytest

import mscomplex_amd as mx
import oracle

pytestmark = pytest.mark.gpu

MB = 1 << 20


def _grid(kind, sx, sy, seed):
    if kind == "g":
        return mx.gen_gaussians(sx, sy, 8, 1e-3, seed=seed)
    return mx.gen_value_noise(sx, sy, seed=seed, period=max(sx, sy) / 4)


def _mix(n, seed):
    r = np.random.default_rng(seed)
    out = []
    for i in range(n):
        sx = int(np.exp(r.uniform(np.log(24), np.log(260))))
        sy = int(np.clip(sx * r.uniform(0.25, 1.0) if r.random() < 0.5 else sx / r.uniform(0.25, 1.0), 8, 400))
        lx = int(r.integers(0, 3)) if min(sx, sy) > 60 else 0
        ly = int(r.integers(0, 3)) if min(sx, sy) > 60 else 0
        th = float(r.choice([0.0, 0.005, 0.01, 0.03, 0.05]))
        out.append((_grid("g" if i % 2 else "v", sx, sy, seed * 100 + i), lx, ly, th))
    return out


def _plain(g, lx, ly, th):
    return mx.GridRun(g, levels=0, lx=lx, ly=ly, threshold=th)


def _same(view, run, label):
    assert view.num_pieces() == run.num_pieces(), label
    for p in range(run.num_pieces()):
        assert np.array_equal(view.piece(p).dump(), run.piece(p).dump()), f"{label}: piece {p} differs"


def test_mix_equals_alone_and_plain():
    reqs = _mix(16, seed=11)
    print(f"[serve] mix of {len(reqs)}", flush=True)
    srv = mx.Server(8 << 30, host_threads=4)
    tickets = [srv.submit(g, lx, ly, th) for g, lx, ly, th in reqs]
    for k, (t, (g, lx, ly, th)) in enumerate(zip(tickets, reqs)):
        st, err, lat, view = srv.wait(t, timeout=120)
        assert st == 0, err
        assert lat > 0
        _same(view, _plain(g, lx, ly, th), f"request {k} in the mix")
        srv.release(t)
    # the same requests one at a time through the same server
    for k, (g, lx, ly, th) in enumerate(reqs[:6]):
        t = srv.submit(g, lx, ly, th)
        st, err, _, view = srv.wait(t, timeout=120)
        assert st == 0, err
        _same(view, _plain(g, lx, ly, th), f"request {k} alone")
        srv.release(t)
    s = srv.stats()
    assert s["done"] == 22 and s["rounds"] >= 7
    srv.close()


def test_oracle_parity_small():
    g = _grid("v", 96, 80, 5)
    srv = mx.Server(1 << 30, host_threads=2)
    t = srv.submit(g, 1, 1, 0.02)
    st, err, _, view = srv.wait(t, timeout=120)
    assert st == 0, err
    ref = oracle.Pipeline(g, lx=1, ly=1, threshold=0.02)
    for p in range(ref.num_pieces()):
        assert np.array_equal(ref.dump(4, p), view.piece(p).dump()), f"piece {p}"
    srv.release(t)
    srv.close()


def test_staggered_arrivals_overlap():
    """Requests arriving while others are in flight join later rounds."""
    reqs = _mix(10, seed=12)
    srv = mx.Server(8 << 30, host_threads=4)
    tickets = []
    for g, lx, ly, th in reqs:
        tickets.append(srv.submit(g, lx, ly, th))
        time.sleep(0.003)
    for t, (g, lx, ly, th) in zip(tickets, reqs):
        st, err, _, view = srv.wait(t, timeout=120)
        assert st == 0, err
        _same(view, _plain(g, lx, ly, th), "staggered")
        srv.release(t)
    srv.close()


def test_repeat_requests_resimplify_resident_grid():
    g = _grid("v", 180, 150, 21)
    srv = mx.Server(2 << 30, host_threads=4)
    t0 = srv.submit(g, 1, 1, 0.01, key=77)
    # a repeat submitted while the first is in flight waits for it
    t1 = srv.submit(g, 1, 1, 0.04, key=77)
    for t, th in ((t0, 0.01), (t1, 0.04)):
        st, err, _, view = srv.wait(t, timeout=120)
        assert st == 0, err
        _same(view, _plain(g, 1, 1, th), f"key 77 at {th}")
    # repeats without the data, served from the resident grid
    for th in (0.0, 0.02, 0.05):
        t = srv.submit(None, 1, 1, th, key=77, size=(g.shape[1], g.shape[0]))
        st, err, _, view = srv.wait(t, timeout=120)
        assert st == 0, err
        _same(view, _plain(g, 1, 1, th), f"resident key 77 at {th}")
        srv.release(t)
    srv.release(t0)
    srv.release(t1)
    assert srv.stats()["repeats"] == 4
    srv.close()


def test_repeats_from_device_resident_complexes(monkeypatch):
    # with the merge tree on the GPU the resident grid keeps its merged
    # complexes in HBM; repeats copy them there, simplify at their own
    # threshold and merge down on the GPU
    monkeypatch.setenv("MSCX_DEVICE_MERGE_MIN_VERTICES", "1")
    g = _grid("v", 300, 260, 23)
    srv = mx.Server(2 << 30, host_threads=4)
    s0 = mx.merge_stats()
    t0 = srv.submit(g, 2, 1, 0.01, key=91)
    st, err, _, view = srv.wait(t0, timeout=120)
    assert st == 0, err
    _same(view, mx.GridRun(g, levels=0, lx=2, ly=1, threshold=0.01, merge="host"), "device-resident first run")
    for th in (0.0, 0.03, 0.05):
        t = srv.submit(None, 2, 1, th, key=91, size=(g.shape[1], g.shape[0]))
        st, err, _, view = srv.wait(t, timeout=120)
        assert st == 0, err
        _same(view, mx.GridRun(g, levels=0, lx=2, ly=1, threshold=th, merge="host"), f"device-resident repeat {th}")
        srv.release(t)
    s1 = mx.merge_stats()
    # 7 merges up once, 7 down per request, no fallback
    assert s1[1] == s0[1] and s1[0] - s0[0] == 7 * 5, (s0, s1)
    srv.release(t0)
    assert srv.stats()["repeats"] == 3
    srv.close()


def test_budget_admission_eviction_and_rejection():
    grids = [_grid("g", 220, 200, 31 + i) for i in range(4)]
    one = 0
    srv0 = mx.Server(1 << 30, host_threads=2)
    t = srv0.submit(grids[0], 0, 0, 0.01, key=1)
    srv0.wait(t, timeout=120)
    one = srv0.stats()["device_bytes"]
    srv0.release(t)
    srv0.close()
    assert one > 0
    # room for about two resident grids: the later keys evict the earlier ones
    srv = mx.Server(int(2.5 * one), host_threads=2)
    ts = [srv.submit(g, 0, 0, 0.01, key=10 + i) for i, g in enumerate(grids)]
    for t, g in zip(ts, grids):
        st, err, _, view = srv.wait(t, timeout=120)
        assert st == 0, err
        _same(view, _plain(g, 0, 0, 0.01), "under a tight budget")
        srv.release(t)
    s = srv.stats()
    assert s["evicted"] >= 1 and s["resident"] <= 3
    # a grid larger than the whole budget is refused on its own
    big = _grid("v", 1200, 1100, 9)
    t = srv.submit(big, 0, 0, 0.0)
    st, err, _, _ = srv.wait(t, timeout=120)
    assert st == -4 and "budget" in err
    srv.release(t)
    srv.close()


def test_failing_request_is_isolated():
    good = _grid("v", 64, 48, 41)
    bad = np.full((20, 20), np.nan, np.float32)
    srv = mx.Server(1 << 30, host_threads=2)
    tg, tb = srv.submit(good, 0, 0, 0.01), srv.submit(bad, 0, 0, 0.01)
    st, err, _, _ = srv.wait(tb, timeout=60)
    assert st == -1 and "NaN" in err
    st, err, _, view = srv.wait(tg, timeout=60)
    assert st == 0, err
    _same(view, _plain(good, 0, 0, 0.01), "good next to a failing request")
    srv.release(tg)
    srv.release(tb)
    srv.close()


def test_concurrent_clients():
    reqs = _mix(12, seed=13)
    srv = mx.Server(8 << 30, host_threads=4)
    results = {}

    def client(k0):
        for k in range(k0, len(reqs), 3):
            g, lx, ly, th = reqs[k]
            t = srv.submit(g, lx, ly, th)
            results[k] = srv.wait(t, timeout=120)[:2] + (t,)

    th_ = [threading.Thread(target=client, args=(i,)) for i in range(3)]
    for x in th_:
        x.start()
    for x in th_:
        x.join()
    for k, (g, lx, ly, th) in enumerate(reqs):
        st, err, t = results[k]
        assert st == 0, err
        _, _, _, view = srv.wait(t, timeout=1)
        _same(view, _plain(g, lx, ly, th), f"client request {k}")
        srv.release(t)
    srv.close()


def test_destroy_with_requests_queued():
    """Destroying the server while requests wait for memory returns: the
    admitted ones finish, the queued ones are failed, nothing hangs."""
    grids = [_grid("v", 300, 300, 50 + i) for i in range(6)]
    one = mx.GridRun(grids[0], levels=0)
    del one
    # a budget that admits about one grid at a time
    srv = mx.Server(24 * MB, host_threads=2)
    for g in grids:
        srv.submit(g, 0, 0, 0.01)
    done = threading.Event()

    def close():
        srv.close()
        done.set()

    th = threading.Thread(target=close, daemon=True)
    th.start()
    assert done.wait(120), "mscx_server_destroy did not return"


def test_borrowed_views_keep_owner_and_expire():
    g = _grid("v", 120, 100, 3)
    ref = mx.GridRun(g, levels=0, lx=1, ly=1, threshold=0.01)
    # a piece of a temporary run stays valid: it keeps the run alive
    d = mx.GridRun(g, levels=0, lx=1, ly=1, threshold=0.01).piece(0).dump()
    assert np.array_equal(d, ref.piece(0).dump())
    srv = mx.Server(1 << 30, host_threads=2)
    t = srv.submit(g, 1, 1, 0.01)
    st, err, _, view = srv.wait(t, timeout=120)
    assert st == 0, err
    piece = view.piece(0)
    assert np.array_equal(piece.dump(), ref.piece(0).dump())
    srv.release(t)
    with pytest.raises(mx.MscxError):
        view.piece(0)
    with pytest.raises(mx.MscxError):
        piece.dump()
    srv.close()


def test_dense_request_fails_fast():
    """A request whose simplification would grow arc multisets without bound
    (white noise at 10%) fails quickly under the server's work limit, and
    the other requests of the mix are unaffected."""
    from tests.helpers import rng_field
    bad = rng_field("white", 150, 130, 5)
    good = _grid("v", 140, 120, 8)
    srv = mx.Server(1 << 30, host_threads=2)
    t_bad = srv.submit(bad, 0, 0, 0.1)
    t_good = srv.submit(good, 1, 1, 0.02)
    st, err, lat, _ = srv.wait(t_bad, timeout=120)
    assert st in (-4, -7), (st, err)
    assert lat < 5.0, lat
    st, err, _, view = srv.wait(t_good, timeout=120)
    assert st == 0, err
    _same(view, _plain(good, 1, 1, 0.02), "request beside a failing one")
    srv.release(t_bad)
    srv.release(t_good)
    srv.close()
