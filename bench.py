#!/usr/bin/env python3
"""Benchmark: Morse-Smale complex of a 4096x4096 float32 value-noise grid on
one MI355X, with a bounded CPU-baseline sample timed on the host.

python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = the end-to-end Morse-Smale complex of the grid, input already
resident in HBM: the 16 leaf tiles of the 4x4 tiling go through gradient,
critical collation, manifold tracing and saddle incidence on the GPU (one
batched launch per stage), then the complexes are merged up the tile tree,
simplified at 1% of the critical-value range, merged down, and every leaf's
manifold discs are filled.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# The rocprofv3 counter summary of the main workload's tile stages: its
# "traffic_bytes_per_launch: N" line is the HBM traffic of one k_grad_local
# launch (FETCH_SIZE and WRITE_SIZE scaled by the factors a device copy of
# known size gives), read at run time so that the figure is the profile's.
COUNTER_PROFILE = os.path.join(ROOT, "profiles", "tile_stages_4096_vnoise_4x4_counters.md")


def profiled_traffic():
    import re
    try:
        with open(COUNTER_PROFILE) as f:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None

SIDE = 4096
LX = LY = 2          # 4x4 tiles
OCTAVES, LACUNARITY, GAIN = 8, 2.0, 0.5
PERIOD = SIDE / 4.0  # base lattice period in pixels
THRESHOLD = 0.01
SEED = 20240601


def split_pieces(sx, sy, lx, ly):
    """Leaf (rect, ext) in piece order: the reference's split/halo rule
    (grid_datamanager.cpp:53-89) applied alternately along y and x."""
    out = []

    def rec(r, e, lx, ly):
        if lx == 0 and ly == 0:
            out.append((r, e))
            return
        dim = 1 if (ly >= lx and ly > 0) else 0
        s = [r[2] - r[0], r[3] - r[1]]
        tr1 = [r[2], r[3]]
        tr1[dim] -= 2 * (s[dim] // 4)
        bl1 = [r[0], r[1]]
        bl1[dim] = tr1[dim]
        tr2 = [e[2], e[3]]
        tr2[dim] = tr1[dim] + 2
        bl2 = [e[0], e[1]]
        bl2[dim] = bl1[dim] - 2
        nlx, nly = (lx, ly - 1) if dim == 1 else (lx - 1, ly)
        rec((r[0], r[1], tr1[0], tr1[1]), (e[0], e[1], tr2[0], tr2[1]), nlx, nly)
        rec((bl1[0], bl1[1], r[2], r[3]), (bl2[0], bl2[1], e[2], e[3]), nlx, nly)

    r = (0, 0, 2 * (sx - 1), 2 * (sy - 1))
    rec(r, r, lx, ly)
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(data):
    """The oracle (the C++ restatement of the reference's CPU path) on the
    workload itself: the full 4096x4096 grid, 4x4 tiles, simplification at
    1%, manifold discs, in the GridDataManager order.  Once with the
    independent pieces of each phase on `threads` host threads (the
    reference's --num-parallel = cores; 16 = this process's CPU share of a
    one-GPU host at most) -- the reported value -- and once single-threaded
    (-st).  Simplification is sequential in both.  Per-phase times are
    reported beside the product's."""
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    names = ("compute", "merge_up", "simplify", "merge_down", "discs")
    out = {}
    for th in (threads, 1):
        t0 = time.perf_counter()
        p = oracle.Pipeline(data, lx=LX, ly=LY, threshold=THRESHOLD, keep=False, threads=th)
        el = time.perf_counter() - t0
        ph = p.times()
        del p
        out[th] = (el, ph)
        print(f"[bench] cpu baseline on {th} threads: {el:.2f}s", file=sys.stderr, flush=True)
    el_mt, ph_mt = out[threads]
    el_st, ph_st = out[1]
    nm = data.size
    return {"value": round(nm / 1e6 / el_mt, 3), "unit": "Mcells/s", "cores": threads,
            "kind": "port",
            "sample": f"the whole workload once: oracle end-to-end pipeline on the {SIDE}x{SIDE} grid "
                      f"(4x4 tiles, simplification at {THRESHOLD}, manifold discs), {el_mt:.2f}s on "
                      f"{threads} threads (--num-parallel form)",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "ms_per_step": round(1e3 * el_mt, 1),
            "phase_ms": dict(zip(names, [round(1e3 * v, 2) for v in ph_mt])),
            "single_thread": {"value": round(nm / 1e6 / el_st, 3), "unit": "Mcells/s",
                              "cores": 1, "ms_per_step": round(1e3 * el_st, 1),
                              "sample": f"the whole workload once, {el_st:.2f}s (-st form)",
                              "phase_ms": dict(zip(names, [round(1e3 * v, 2) for v in ph_st]))}}


def worstcase_bench(reps=3):
    """The worst-case configuration's GPU tile stages: 16384x16384 white
    noise U[0,1) and its 256-level quantisation, 8x8 tiling, inputs
    resident in HBM; every leaf tile through gradient, collation, trace and
    incidence in one launch per stage (mscx_tiles_work).  The host merge
    tree at ~2e8 critical points is not run here (its host phases alone
    would take many minutes)."""
    import mscomplex_amd as mx
    side = 16384
    out = {}
    for levels in (0, 256):
        data = mx.gen_white(side, side, levels=levels, seed=4242 + levels)
        buf = mx.DeviceBuffer.from_array(data)
        pieces = split_pieces(side, side, 3, 3)
        tiles = []
        for rect, ext in pieces:
            t = mx.GridDataset(rect, ext)
            t.bind_device(buf.ptr + 4 * ((ext[1] // 2) * side + ext[0] // 2), side)
            tiles.append(t)
        mx.work_tiles(tiles)
        mx.profile(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            mx.work_tiles(tiles)
        el = (time.perf_counter() - t0) / reps
        prof = mx.profile_read()
        mx.profile(False)
        K = sum(t.num_critical() for t in tiles)
        geo = [((e[2] - e[0]) // 2 + 1, (e[3] - e[1]) // 2 + 1) for _, e in pieces]
        nm = sum(vx * vy for vx, vy in geo)
        cells = sum((2 * vx - 1) * (2 * vy - 1) for vx, vy in geo)
        per = {k: v[0] / max(v[1], 1) for k, v in prof.items()}
        g_ms = per["gradient"]
        gt_ms = per["gradient"] + per["trace"] + per["owner"]
        gt_bytes = 4 * nm + 6 * cells
        out["white" if levels == 0 else f"quantized{levels}"] = {
            "critical_cells": int(K), "tile_stages_ms": round(1e3 * el, 3),
            "tile_stages_Mcells_per_s": round(side * side / 1e6 / el, 1),
            "stage_ms": {k: round(v, 4) for k, v in per.items()},
            "k_grad_local_GBs": round((4 * nm + cells) / (g_ms * 1e-3) / 1e9, 1),
            "gradient_trace_GBs": round(gt_bytes / (gt_ms * 1e-3) / 1e9, 1),
            "gradient_trace_frac": round(gt_bytes / (gt_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        print(f"[bench] worst case levels={levels}: {1e3 * el:.2f} ms, K={K}", file=sys.stderr, flush=True)
        del tiles, buf, data
    return out


def worst_e2e_child(side, levels, threshold):
    """One worst-case grid end to end (child process, so that its peak host
    memory is its own): prints one JSON line."""
    import resource
    import mscomplex_amd as mx
    data = mx.gen_white(side, side, levels=levels, seed=4242 + levels)
    buf = mx.DeviceBuffer.from_array(data)
    del data
    s0 = mx.merge_stats()
    t0 = time.perf_counter()
    run = mx.GridRun(None, lx=3, ly=3, threshold=threshold, device_ptr=buf.ptr, size=(side, side))
    el = time.perf_counter() - t0
    s1 = mx.merge_stats()
    n = run.num_pieces()
    print(json.dumps({"side": side, "levels": levels, "threshold": threshold,
                      "merge_tree": "device (HBM-resident complexes)" if s1[0] > s0[0] and s1[1] == s0[1] else "host",
                      "end_to_end_s": round(el, 2), "Mcells_per_s": round(side * side / 1e6 / el, 2),
                      "phase_s": dict(zip(("compute", "merge_up", "simplify", "merge_down", "discs"),
                                          [round(v, 2) for v in run.times()])),
                      "root_critical_points": run.piece_num_cps(n - 1),
                      "peak_host_rss_GB": round(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1e6, 1)}),
          flush=True)


def worst_e2e_bench(side=16384, threshold=0.0):
    """The worst-case configuration end to end at full size: white noise and
    its 256-level quantisation, 8x8 tiles, the merge tree on the GPU, each
    in a child process (its peak host memory is its own)."""
    import subprocess
    out = {}
    for levels in (0, 256):
        if levels:
            # the previous child's ~45 GB of host memory (and its device
            # allocations) are still being returned for a few seconds after it
            # exits; a child started at once measured its merge up 2.6-4.3 s
            # instead of 0.36 s
            time.sleep(10)
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--worst-e2e-child", str(side), str(levels),
                            str(threshold)], capture_output=True, text=True, timeout=1500)
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
        out["white" if levels == 0 else f"quantized{levels}"] = json.loads(line[-1]) if line else \
            {"error": r.stderr[-2000:], "returncode": r.returncode}
        print(f"[bench] worst case e2e levels={levels}: {out}", file=sys.stderr, flush=True)
    return out


def batched_bench(n=1024, seed=7):
    """The batched configuration: n grids with long sides log-uniform in
    [64, 2048], aspect 1:1-1:4, half Gaussian sums and half value noise,
    threshold 1%, inputs resident in HBM.  Batched = every grid's GPU stages
    in one launch per stage and the host phases of all grids in parallel;
    plain = one mscx_grid_run per grid.  Leaf tilings per size class come
    from the autotuner (timed on a 64-grid sample, outside the timed runs).
    The batched figure is the steady state: one untimed batch first, then
    the mean of two."""
    import mscomplex_amd as mx
    t0 = time.perf_counter()
    mix = mx.uneven_mix(n, 64, 2048, seed, aspect_max=4.0, thresholds=0.01)
    grids = [g for g, _ in mix]
    # Gaussian sums (odd indices) at threshold 0: their +/-1e-3 noise makes
    # the reference's simplification multiply strangulated arcs without
    # bound at 1% (it fails only at the arc-entry cap, after minutes)
    ths = [0.0 if i % 2 else t for i, (_, t) in enumerate(mix)]
    t_gen = time.perf_counter() - t0
    print(f"[bench] batched: generated {n} grids in {t_gen:.1f}s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sample = mx.uneven_mix(64, 64, 2048, seed + 1, aspect_max=4.0, thresholds=0.01)
    sample_ths = [0.0 if i % 2 else t for i, (_, t) in enumerate(sample)]
    policy, table = mx.autotune_tiling([g for g, _ in sample], sample_ths)
    # kernel block height per tile-height class, timed on the GPU stages
    rows_table, rows_meas = mx.autotune_block_rows([g for g, _ in sample], sample_ths, tiling_policy=policy)
    t_tune = time.perf_counter() - t0
    tiling = [policy(g.shape[1], g.shape[0]) for g in grids]
    bufs = [mx.DeviceBuffer.from_array(g) for g in grids]
    dev = [(b.ptr, g.shape[1], g.shape[0]) for b, g in zip(bufs, grids)]
    lx = [t[0] for t in tiling]
    ly = [t[1] for t in tiling]
    # one untimed batch warms the library's caches to the batch's size
    # (device pool blocks, pinned staging, manifold-pass leases); then the
    # mean of two timed batches
    t0 = time.perf_counter()
    del_b = mx.Batch(dev, lx, ly, ths, device=True)
    t_first = time.perf_counter() - t0
    del del_b
    t_batch = 0.0
    for rep in range(2):
        if rep:
            del b
        t0 = time.perf_counter()
        b = mx.Batch(dev, lx, ly, ths, device=True)
        t_batch += (time.perf_counter() - t0) / 2
    print(f"[bench] batched: autotune {t_tune:.1f}s, first batch {t_first:.2f}s, batch {t_batch:.2f}s",
          file=sys.stderr, flush=True)
    bad = [i for i in range(n) if b.status(i)[0] != 0]
    # word-for-word parity of every grid, batched vs plain (checksums of the
    # pieces' dump words, outside the timed regions)
    sums_b = [b.grid(i).checksum() if i not in bad else None for i in range(n)]
    del b
    t_plain = 0.0
    mismatch = []
    for i, ((p, sx, sy), a, c, th) in enumerate(zip(dev, lx, ly, ths)):
        t0 = time.perf_counter()
        r = mx.GridRun(None, levels=0, lx=a, ly=c, threshold=th, device_ptr=p, size=(sx, sy))
        t_plain += time.perf_counter() - t0
        if sums_b[i] is not None and r.checksum() != sums_b[i]:
            mismatch.append(i)
        del r
    print(f"[bench] batched: plain {t_plain:.2f}s, {len(mismatch)} grids differ from the batch", file=sys.stderr,
          flush=True)
    vert = sum(g.size for g in grids)
    return {"grids": n, "vertices": int(vert), "batched_grids_per_s": round(n / t_batch, 2),
            "plain_grids_per_s": round(n / t_plain, 2), "speedup": round(t_plain / t_batch, 2),
            "batched_s": round(t_batch, 3), "plain_s": round(t_plain, 3), "failed": bad,
            "timing": "mean of two batches after one untimed warm-up batch; first (cold) batch "
                      f"{round(t_first, 3)} s",
            "parity_vs_plain": {"grids_compared": n - len(bad), "differing": mismatch,
                                "method": "FNV-1a of every piece's dump words (mscx_grid_checksum)"},
            "tiling_policy": {"side<256": policy(255, 255), "side<1024": policy(1023, 1023),
                              "side>=1024": policy(2048, 2048)},
            "autotune_table": {f"{b}:{c}": round(v, 4) for (b, c), v in table.items()},
            "block_rows_policy": {"tile_rows<=" + str(b): r for b, r in rows_table} or "32 rows for every class",
            "block_rows_gpu_stage_ms": {f"tile_rows<{b}:{r}": round(v, 3) for (b, r), v in rows_meas.items()},
            "autotune_s": round(t_tune, 2), "generate_s": round(t_gen, 2)}


def served_bench(n=100, seed=11, repeat_frac=0.25, loads=((2.0, 5), (8.0, 6)), max_side=8192, work_limit=128):
    """The served configuration: Poisson arrivals of square grids with sides
    log-uniform in [128, max_side], alternating Gaussian sums (32 Gaussians,
    +/-1e-3 noise) and value noise, every request at a threshold uniform in
    [0, 5%] of its range, a quarter of the requests repeating an earlier grid
    at a new threshold.  Plain = mscx_grid_run per request, one at a time
    (repeats recompute).  Served = arrivals at each of `loads` (x the plain
    throughput) into the continuous-batching server (200 GB HBM budget,
    grids resident under their key for the repeats), each run starting with
    an empty server.  Both forms run the simplification under the same work
    limit (x its initial heap entries): a noisy grid at a large threshold,
    which the reference's exact simplification would grind through for
    minutes, fails in seconds in both and counts as a failed request; its
    latency is included in the percentiles of all requests."""
    import threading
    import mscomplex_amd as mx
    r = np.random.default_rng(seed)
    base = []
    reqs = []  # (grid index, threshold)
    for i in range(n):
        if base and r.random() < repeat_frac:
            reqs.append((int(r.integers(len(base))), float(r.uniform(0, 0.05))))
            continue
        side = int(round(np.exp(r.uniform(np.log(128), np.log(max_side)))))
        s = int(r.integers(1 << 62))
        g = mx.gen_gaussians(side, side, 32, 1e-3, seed=s) if len(base) % 2 else \
            mx.gen_value_noise(side, side, seed=s, period=side / 4)
        base.append(g)
        reqs.append((len(base) - 1, float(r.uniform(0, 0.05))))
    tiling = [(0, 0) if max(g.shape) < 1024 else (1, 1) if max(g.shape) < 4096 else (2, 2) for g in base]
    pct = lambda v, q: round(float(np.percentile(v, q)) * 1e3, 2) if len(v) else None

    os.environ["MSCX_WORK_LIMIT"] = str(work_limit)
    plain_lat, plain_ok = [], []
    try:
        t0 = time.perf_counter()
        for k, (gi, th) in enumerate(reqs):
            a = time.perf_counter()
            lx, ly = tiling[gi]
            try:
                run = mx.GridRun(base[gi], levels=0, lx=lx, ly=ly, threshold=th)
                del run
                plain_ok.append(True)
            except mx.MscxError:
                plain_ok.append(False)
            plain_lat.append(time.perf_counter() - a)
            print(f"[bench] served: plain {k + 1}/{n}: side {base[gi].shape[0]}, threshold {th:.4f}, "
                  f"{plain_lat[-1]:.2f}s{'' if plain_ok[-1] else ' (failed)'}", file=sys.stderr, flush=True)
        t_plain = time.perf_counter() - t0
    finally:
        del os.environ["MSCX_WORK_LIMIT"]
    print(f"[bench] served: plain {t_plain:.2f}s for {n} requests", file=sys.stderr, flush=True)

    def run_at(load, gap_seed):
        srv = mx.Server(200 << 30)
        srv.set_work_limit(work_limit)
        rate = load * n / t_plain
        gaps = np.random.default_rng(gap_seed).exponential(1.0 / rate, size=n)
        tickets = [None] * n

        def client():
            t_next = time.perf_counter()
            for k, (gi, th) in enumerate(reqs):
                t_next += gaps[k]
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                lx, ly = tiling[gi]
                tickets[k] = srv.submit(base[gi], lx, ly, th, key=gi + 1)

        s0 = srv.stats()
        t0 = time.perf_counter()
        th_ = threading.Thread(target=client)
        th_.start()
        th_.join()
        lat, ok_lat, codes = [], [], {}
        for k, t in enumerate(tickets):
            st, err, l, _ = srv.wait(t, timeout=900)
            lat.append(l)
            if st == 0:
                ok_lat.append(l)
            else:
                codes[str(st)] = codes.get(str(st), 0) + 1
            srv.release(t)
            print(f"[bench] served: load {load}x, request {k + 1}/{n}: status {st}, {l:.2f}s", file=sys.stderr,
                  flush=True)
        t_served = time.perf_counter() - t0
        s1 = srv.stats()
        srv.close()
        print(f"[bench] served: load {load}x plain, {t_served:.2f}s", file=sys.stderr, flush=True)
        return {"load_x_plain": load, "offered_rate_per_s": round(rate, 3),
                "throughput_per_s": round(n / t_served, 3), "x_plain": round(t_plain / t_served, 2),
                "latency_ms_all": {"p50": pct(lat, 50), "p95": pct(lat, 95), "p99": pct(lat, 99)},
                "latency_ms_completed": {"p50": pct(ok_lat, 50), "p95": pct(ok_lat, 95), "p99": pct(ok_lat, 99)},
                "completed": len(ok_lat), "failed_by_status": codes,
                "rounds": s1["rounds"] - s0["rounds"], "repeats_served_resident": s1["repeats"] - s0["repeats"]}

    runs = [run_at(load, gs) for load, gs in loads]
    ok_lat = [l for l, ok in zip(plain_lat, plain_ok) if ok]
    return {"requests": n, "distinct_grids": len(base), "repeats": n - len(base), "max_side": max_side,
            "sides": {"min": int(min(g.shape[0] for g in base)), "max": int(max(g.shape[0] for g in base))},
            "thresholds": "uniform [0, 0.05] for every request", "work_limit": work_limit,
            "vertices_mean": int(np.mean([base[gi].size for gi, _ in reqs])),
            "plain_throughput_per_s": round(n / t_plain, 3), "plain_s": round(t_plain, 2),
            "plain_completed": int(sum(plain_ok)),
            "plain_latency_ms_all": {"p50": pct(plain_lat, 50), "p95": pct(plain_lat, 95), "p99": pct(plain_lat, 99)},
            "plain_latency_ms_completed": {"p50": pct(ok_lat, 50), "p95": pct(ok_lat, 95), "p99": pct(ok_lat, 99)},
            "served": runs, "hbm_budget_GB": 200}


def copy_probe():
    """A plain copy of known size: the counters' bytes for it calibrate the
    HBM traffic figures read from rocprofv3 (FETCH_SIZE, WRITE_SIZE)."""
    import torch
    n = 64 * 1024 * 1024  # 256 MiB of f32
    a = torch.rand(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy probe: 3 copies of {4 * n} B", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-worst", action="store_true")
    ap.add_argument("--served-max-side", type=int, default=8192)
    ap.add_argument("--served-requests", type=int, default=100)
    ap.add_argument("--batched-grids", type=int, default=1024)
    ap.add_argument("--no-worst-e2e", action="store_true",
                    help="skip the 16384^2 worst case end to end (two grids, ~20 s each with the merge tree on "
                         "the GPU)")
    ap.add_argument("--worst-e2e-child", nargs=3, metavar=("SIDE", "LEVELS", "THRESHOLD"), help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run three 256 MiB device-to-device copies (the calibration kernel for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.worst_e2e_child:
        worst_e2e_child(int(args.worst_e2e_child[0]), int(args.worst_e2e_child[1]), float(args.worst_e2e_child[2]))
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path is single-GPU "
              f"in this build and multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)

    import mscomplex_amd as mx

    data = mx.gen_value_noise(SIDE, SIDE, OCTAVES, LACUNARITY, GAIN, PERIOD, SEED)
    dev = mx.DeviceBuffer.from_array(data)
    pieces = split_pieces(SIDE, SIDE, LX, LY)

    def step():
        return mx.GridRun(None, lx=LX, ly=LY, threshold=THRESHOLD, device_ptr=dev.ptr, size=(SIDE, SIDE))

    for _ in range(args.warmup):
        step()
    # stage timing: HIP events on the work stream, recorded by the replayed
    # stage graph itself (the default submission path) and around the
    # directly launched compaction and incidence kernels
    mx.profile(True)
    g0 = mx.graph_stats()
    host_t = np.zeros(5)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run = step()
        host_t += run.times()
        del run
    el = time.perf_counter() - t0
    prof = mx.profile_read()
    g1 = mx.graph_stats()
    mx.profile(False)
    last = step()
    K = sum(last.piece(i).num_cps() for i in range(len(pieces)))
    root_cps = last.piece(last.num_pieces() - 1).num_cps()

    ms_step = 1000.0 * el / args.steps
    mcells = SIDE * SIDE / 1e6 / (el / args.steps)
    ncells = (2 * SIDE - 1) * (2 * SIDE - 1)

    # the same step from a host array: the upload of the grid is inside the
    # timed region (results stay on the host side of the library either way)
    host_steps = max(2, min(5, args.steps))
    t0 = time.perf_counter()
    for _ in range(host_steps):
        run = mx.GridRun(data, lx=LX, ly=LY, threshold=THRESHOLD)
        del run
    el_host = (time.perf_counter() - t0) / host_steps

    # the same step from a raw grid file (the reference's -gf input, page
    # cache warm): the file is streamed into HBM inside the timed region;
    # the discs are computed but not written out as files
    fd, fpath = tempfile.mkstemp(suffix=".raw")
    os.close(fd)
    try:
        data.tofile(fpath)
        mx.GridRun.from_file(fpath, (SIDE, SIDE), lx=LX, ly=LY, threshold=THRESHOLD)
        file_c = 0.0
        t0 = time.perf_counter()
        for _ in range(host_steps):
            run = mx.GridRun.from_file(fpath, (SIDE, SIDE), lx=LX, ly=LY, threshold=THRESHOLD)
            file_c += run.times()[0]
            del run
        el_file = (time.perf_counter() - t0) / host_steps
    finally:
        os.unlink(fpath)

    # algorithmic bytes per launch (all 16 tiles, halos included), per unit:
    #   gradient (k_grad_local): 4 B per vertex in (f32) + 1 B per cell out (code)
    #   trace (ring resolution + k_owner): 1 B per cell in (code) + 4 B per cell out (owner)
    geo = [((e[2] - e[0]) // 2 + 1, (e[3] - e[1]) // 2 + 1) for _, e in pieces]
    nm = sum(vx * vy for vx, vy in geo)
    cells = sum((2 * vx - 1) * (2 * vy - 1) for vx, vy in geo)
    g_bytes = 4 * nm + cells
    t_bytes = cells + 4 * cells
    per = {k: v[0] / max(v[1], 1) for k, v in prof.items()}  # ms per launch
    stage_ms = {k: v[0] / args.steps for k, v in prof.items()}
    g_ms = per["gradient"]
    gt_ms = per["gradient"] + per["trace"] + per["owner"]
    ach = g_bytes / (g_ms * 1e-3) / 1e9
    gt_ach = (g_bytes + t_bytes) / (gt_ms * 1e-3) / 1e9
    line = {
        "metric": "Mcells/s end-to-end MS complex @4096^2 f32",
        "value": round(mcells, 2),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "4096x4096 f32 value noise (8 octaves, lacunarity 2, gain 0.5, period 1024), "
                               "4x4 tile merge, persistence simplification at 1% of range, manifold discs",
                   "grid": [SIDE, SIDE], "tiles": [4, 4], "threshold": THRESHOLD,
                   "leaf_critical_points": K, "root_critical_points": root_cps},
        "roofline": {"bound": "hbm", "kernel": "k_grad_local", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(),
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated, read from "
                                       + os.path.relpath(COUNTER_PROFILE, ROOT),
                     "timing": "HIP events on the work stream inside the replayed stage graph, "
                               f"{g1[1] - g0[1]} graph replays in the timed steps",
                     "algorithmic_bytes": g_bytes, "avg_launch_ms": round(g_ms, 4),
                     "limiter": "VALU issue: the gradient rules' integer compares and selects "
                                "(~124 VALU per interior vertex row step of a lane, ~1830 per wave with "
                                "the in-block pointer jumping), not HBM bandwidth"},
        "gradient_trace": {"kernels": ["k_grad_local", "k_ring_resolve", "k_owner"],
                           "algorithmic_bytes": g_bytes + t_bytes, "ms": round(gt_ms, 4),
                           "achieved_GBs": round(gt_ach, 1), "frac": round(gt_ach / HBM_PEAK_GBS, 4)},
        "Ccells_per_s_M": round(ncells / 1e6 / (el / args.steps), 2),
        "host_input": {"value": round(SIDE * SIDE / 1e6 / el_host, 2), "unit": "Mcells/s",
                       "ms_per_step": round(1e3 * el_host, 3), "steps": host_steps,
                       "note": "grid handed over in host memory: H2D upload inside the timed region"},
        "file_input": {"value": round(SIDE * SIDE / 1e6 / el_file, 2), "unit": "Mcells/s",
                       "ms_per_step": round(1e3 * el_file, 3), "steps": host_steps,
                       "read_compute_ms": round(1e3 * file_c / host_steps, 3),
                       "resident_compute_ms": round(1e3 * host_t[0] / args.steps, 3),
                       "note": "mscx_grid_run_file on a 64 MiB raw f32 file in the page cache: read + H2D "
                               "in 8 MiB pinned slabs inside the timed region; disc files not written"},
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "phase_ms": dict(zip(("compute", "merge_up", "simplify", "merge_down", "discs"),
                             [round(1000 * v / args.steps, 3) for v in host_t])),
    }
    del last
    if not args.no_worst:
        line["worst_case"] = worstcase_bench()
    if not args.no_worst_e2e:
        line["worst_case_end_to_end"] = worst_e2e_bench()
    if not args.no_batched:
        line["batched"] = batched_bench(n=args.batched_grids)
    if not args.no_served:
        line["served"] = served_bench(n=args.served_requests, max_side=args.served_max_side)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(data)
        line["cpu_baseline"]["product_phase_ms"] = line["phase_ms"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
