"""Benchmark of the discrete-Sibson tile engine on one MI355X.

Workload (one step = one tile through the whole device pipeline, inputs
already resident in HBM): a single 4096 x 4096 tile of 0.5 m cells at
scale 1 with 4M clustered points (2k Matern parents, 30 % water, 1 % exact
duplicates), the plain one-tile-per-launch form.

Prints ONE JSON line with cells/s (value), points/s, the tile latency
percentiles, the roofline of the dominant kernel (k_tiles, HIP events on
its own stream) and the CPU oracle timed on this host's cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def compact(x, sig: int = 4):
    """x with every float rounded to `sig` significant digits (the printed
    line stays short enough for a log tail to hold it whole)."""
    if isinstance(x, float):
        return float(f"{x:.{sig}g}") if np.isfinite(x) else x
    if isinstance(x, dict):
        return {k: compact(v, sig) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [compact(v, sig) for v in x]
    return x


def write_full_record(out: dict, name: str, d: str):
    """The whole record, unrounded and with the per-class autotune tables,
    to <d>/<name>."""
    try:
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, name), "w") as f:
            json.dump(out, f, indent=1)
    except OSError as e:
        log(f"could not write the full record: {e}")


def cpu_baseline(tile, budget_s: float = 12.0):
    """The CPU oracle (site-major raster + query gather, OpenMP) on the same
    tile, repeated until about budget_s of work; returns (cells/s, threads,
    seconds per tile, oracle output of the last run)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    threads = len(os.sched_getaffinity(0))
    threads = min(threads, 16)
    reps, t_total, ref = 0, 0.0, None
    while reps < 1 or (t_total < budget_s and reps < 5):
        t0 = time.perf_counter()
        ref = O.run(tile.params, tile.X, tile.Y, tile.Z, nthreads=threads)
        t_total += time.perf_counter() - t0
        reps += 1
    per = t_total / reps
    return tile.cells / per, threads, per, reps, ref


def cpu_single_thread(tile):
    """The same oracle on one thread, one run: cells/s."""
    import pyoracle as O
    t0 = time.perf_counter()
    O.run(tile.params, tile.X, tile.Y, tile.Z, nthreads=1)
    return tile.cells / (time.perf_counter() - t0)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_bandwidth(ctx, mib: int = 1024, reps: int = 10):
    """STREAM-like device copy (hipMemcpy device to device, one GPU): read +
    write bytes over the best of reps timed copies, GB/s."""
    nb = mib << 20
    a, c = ctx.alloc((nb,), np.uint8), ctx.alloc((nb,), np.uint8)
    ctx.copy(c, a)
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.copy(c, a)  # ends in a stream sync
        best = min(best, time.perf_counter() - t0)
    a.free()
    c.free()
    return 2.0 * nb / best / 1e9


def bench_batch(T, S, ctx, reps: int = 3, tune: bool = True, cpu: bool = True):
    """The 256 uneven tiles (log-normal point counts, grids 256^2..2048^2,
    half uniform, half clustered): the plain form (one tile per launch
    sequence, synchronised) against the batched form (one bucket sort + one
    persistent launch over all tiles, replayed from a captured graph), each
    with the built-in sizes and with sizes the autotuner measured for its
    form on this workload."""
    t0 = time.perf_counter()
    tiles = S.workload_batch()
    log(f"batched workload: {len(tiles)} tiles, {sum(t.n for t in tiles)} points, "
        f"{sum(t.cells for t in tiles)} cells, generated in {time.perf_counter() - t0:.1f}s")
    dev = [(ctx.upload(t.X), ctx.upload(t.Y), ctx.upload(t.Z),
            ctx.alloc((t.params.nrows, t.params.ncols), np.float32)) for t in tiles]
    cells = sum(t.cells for t in tiles)
    pts = sum(t.n for t in tiles)
    P = [t.params for t in tiles]
    cols = ([d[0] for d in dev], [d[1] for d in dev], [d[2] for d in dev])
    ns = [t.n for t in tiles]
    hs = [d[3] for d in dev]

    # the C arguments made once, outside the timed loops (the engine's
    # time, not the Python marshalling of 256 descriptors)
    D = [p.to_desc() for p in P]
    B = ctx.batch_args(P, *cols, ns, hs)

    def plain():
        for t, d, dd in zip(tiles[:8], dev[:8], D[:8]):
            ctx.tile_run_dev(dd, d[0], d[1], d[2], t.n, d[3])
        lat = []
        s0 = time.perf_counter()
        for t, d, dd in zip(tiles, dev, D):
            a = time.perf_counter()
            ctx.tile_run_dev(dd, d[0], d[1], d[2], t.n, d[3])
            lat.append(time.perf_counter() - a)
        return time.perf_counter() - s0, lat, [h.to_host() for h in hs]

    def batched():
        ctx.batch_run_args(B)
        s0 = time.perf_counter()
        for _ in range(reps):
            ctx.batch_run_args(B)
        return (time.perf_counter() - s0) / reps, [h.to_host() for h in hs]

    def same(a, b):
        return all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))

    plain_s, lat, plain_out = plain()
    batch_s, batch_out = batched()
    tuned_b = tuned_p = []
    tune_s = 0.0
    if tune:
        t_tune = time.perf_counter()
        tuned_b = ctx.autotune(P, *cols, ns, batched=True, reps=2)
        tuned_p = ctx.autotune(P, *cols, ns, batched=False, reps=1)
        tune_s = time.perf_counter() - t_tune
        log(f"autotuned {len(tuned_b)} batched + {len(tuned_p)} plain shape classes in {tune_s:.1f}s")
    plain_t_s, lat_t, plain_t_out = plain()
    batch_t_s, batch_t_out = batched()
    for d in dev:
        for a in d:
            a.free()
    # the CPU oracle over every tile, timed (the CPU baseline of this
    # workload in full), and every batched tile against its exact heights
    cpu_base = None
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        threads = min(len(os.sched_getaffinity(0)), 16)
        o_s, bits = 0.0, 0
        for t, h in zip(tiles, batch_out):
            a = time.perf_counter()
            ref = O.run(t.params, t.X, t.Y, t.Z, nthreads=threads)
            o_s += time.perf_counter() - a
            bits += int(np.count_nonzero(h.view(np.uint32) != ref["h_fixed"].view(np.uint32)))
        cpu_base = {"value": cells / o_s, "unit": "cells/s", "cores": threads, "kind": "port",
                    "sample": f"all {len(tiles)} tiles, {o_s:.1f} s",
                    "batched_bits_differ_from_oracle": bits}
        log(f"CPU oracle over the {len(tiles)} tiles: {o_s:.1f}s, {bits} height bits differ from the batched form")

    def pct(v):
        return {"p50": float(np.percentile(v, 50) * 1e3), "p99": float(np.percentile(v, 99) * 1e3)}

    choices = {"batched": [{k: e[k] for k in ("shape_class", "bucket_px", "tile_cells", "wg_per_cu",
                                              "ms_default", "ms_tuned")} for e in tuned_b],
               "plain": [{k: e[k] for k in ("shape_class", "bucket_px", "tile_cells", "wg_per_cu",
                                            "ms_default", "ms_tuned")} for e in tuned_p]}
    return {
        "workload": "256 uneven tiles, log-normal points (median 200k, 10k-2M), grids "
                    "256^2-2048^2 at 0.5 m, half uniform / half clustered",
        "tiles": len(tiles), "points": pts, "cells": cells,
        "plain_cells_per_s": cells / plain_s, "plain_points_per_s": pts / plain_s,
        "plain_tile_ms": pct(lat),
        "batched_cells_per_s": cells / batch_s, "batched_points_per_s": pts / batch_s,
        "batched_ms": batch_s * 1e3,
        "speedup": plain_s / batch_s,
        "tuned": {"classes_batched": len(tuned_b), "classes_plain": len(tuned_p),
                  "tuning_s": tune_s,
                  "plain_cells_per_s": cells / plain_t_s, "plain_tile_ms": pct(lat_t),
                  "batched_cells_per_s": cells / batch_t_s, "batched_ms": batch_t_s * 1e3,
                  "speedup": plain_t_s / batch_t_s} if tune else None,
        "batched_equals_plain": bool(same(plain_out, batch_out) and same(plain_out, plain_t_out)
                                     and same(plain_out, batch_t_out)),
        "cpu_baseline": cpu_base,
    }, tiles, ctx.tune_table(), choices


def bench_uniform_100k(T, S, ctx, steps: int = 20, cpu: bool = True):
    """The reference's own CPU-runnable case: 100k uniform points in 1000 m,
    a 512^2 DEM at 2 m, scale 5 (2584^2 pixels), plain form on device
    buffers; the CPU oracle on the same tile and the height bits that
    differ from it."""
    t = S.workload_uniform_100k()
    dX, dY, dZ = ctx.upload(t.X), ctx.upload(t.Y), ctx.upload(t.Z)
    dH = ctx.alloc((t.params.nrows, t.params.ncols), np.float32)
    for _ in range(3):
        ctx.tile_run_dev(t.params, dX, dY, dZ, t.n, dH)
    s0 = time.perf_counter()
    for _ in range(steps):
        ctx.tile_run_dev(t.params, dX, dY, dZ, t.n, dH)
    g_s = (time.perf_counter() - s0) / steps
    h = dH.to_host()
    for a in (dX, dY, dZ, dH):
        a.free()
    out = {"workload": "100k uniform points, 512x512 DEM at 2 m, scale 5", "points": t.n,
           "cells": t.cells, "gpu_ms": g_s * 1e3, "gpu_cells_per_s": t.cells / g_s}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        threads = min(len(os.sched_getaffinity(0)), 16)
        s0 = time.perf_counter()
        ref = O.run(t.params, t.X, t.Y, t.Z, nthreads=threads)
        o_s = time.perf_counter() - s0
        out["cpu_baseline"] = {"value": t.cells / o_s, "unit": "cells/s", "cores": threads, "kind": "port",
                               "sample": f"the same tile, {o_s:.2f} s"}
        out["bits_differ_from_oracle"] = int(np.count_nonzero(h.view(np.uint32) != ref["h_fixed"].view(np.uint32)))
    log(f"uniform 100k: {out}")
    return out


def bench_served(T, tiles, rates, n_req, resident: bool, seed=4, tune_table=None):
    """Served form: Poisson arrivals at each rate of tiles drawn from the
    uneven-tile pool; one thread submits, another waits in order.  resident:
    the tiles' points and the output rasters live in HBM (tnni_submit_dev);
    otherwise host buffers in and out (tnni_submit).  Either way the caller
    keeps a ring of 512 output rasters and reuses them, as a server would.
    Latency is submit -> results in device memory (resident) or in the
    caller's host buffers, timed by the engine."""
    import threading
    out = []
    rng = np.random.default_rng(seed)
    ring = 512  # output rasters in flight
    # admission by HBM footprint within a 200 GB budget; batches of up to 256
    # tiles (the engine's default), three in flight
    with T.Context(0, n_slots=3, max_batch_tiles=256, hbm_budget=200 << 30) as c:
        if tune_table:
            c.tune_load(tune_table)
        big = max(t.cells for t in tiles)
        if resident:
            dev = [(c.upload(t.X), c.upload(t.Y), c.upload(t.Z)) for t in tiles]
            outs = [c.alloc((big,), np.float32) for _ in range(ring)]
        else:
            outs = [np.zeros(big, np.float32) for _ in range(ring)]

        def submit(k, i):
            t = tiles[i]
            if resident:
                d = dev[i]
                return c.submit_dev(t.params, d[0], d[1], d[2], t.n, outs[k % ring])
            return c.submit(t.params, t.X, t.Y, t.Z, out=outs[k % ring])

        # warm the slots, graphs and (host buffers) the pinned staging pool:
        # one pass of ring requests in flight at once
        warm = [submit(k, k % len(tiles)) for k in range(16 if resident else ring)]
        for p in warm:
            p.wait()
        for rate in rates:
            c.serve_stats_reset()
            # saturation: enough requests that filling and draining the
            # pipeline do not dominate the span
            per_rate = n_req if rate > 0 else max(n_req, 1000)
            picks = rng.integers(0, len(tiles), per_rate)
            # rate 0: saturation, every request submitted as soon as it can be
            gaps = rng.exponential(1.0 / rate, per_rate) if rate > 0 else np.zeros(per_rate)
            pend, lat = [], []
            cv = threading.Condition()
            state = {"waited": 0}

            def waiter():
                for k in range(per_rate):
                    with cv:
                        while len(pend) <= k:
                            cv.wait()
                        p = pend[k]
                    p.wait()
                    lat.append(p.latency_s)
                    with cv:
                        state["waited"] = k + 1
                        cv.notify_all()
                state["t_end"] = time.perf_counter()

            th = threading.Thread(target=waiter)
            th.start()
            t0 = time.perf_counter()
            t_next = t0
            for k in range(per_rate):
                t_next += gaps[k]
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                if resident:
                    with cv:  # the output raster k % ring is free again
                        while state["waited"] < k - ring + 1:
                            cv.wait()
                p = submit(k, int(picks[k]))
                with cv:
                    pend.append(p)
                    cv.notify_all()
            th.join()
            span = state["t_end"] - t0
            cells = float(sum(tiles[i].cells for i in picks))
            st = c.serve_stats()
            p99 = float(np.percentile(lat, 99))
            out.append({"rate": rate, "requests": per_rate,
                        "tiles_per_s": per_rate / span, "cells_per_s": cells / span,
                        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": p99 * 1e3,
                        "p99_tail_n": int(np.count_nonzero(np.asarray(lat) >= p99)),
                        "max_ms": float(np.max(lat) * 1e3),
                        "inflight_tiles": st["max_inflight_tiles"],
                        "inflight_batches": st["max_inflight_batches"]})
            log(f"served ({'HBM-resident' if resident else 'host buffers'}) rate {rate}/s: {out[-1]}")
        if resident:
            for d in dev:
                for a in d:
                    a.free()
            for o in outs:
                o.free()
    return out


def bench_ref100k(T, S, args):
    """--workload ref100k: only the reference's CPU-runnable case (100k
    uniform points in 1000 m, 512^2 DEM at 2 m, scale 5: 2584^2 pixels),
    plain form on device buffers, K steps timed, k_tiles timed with HIP
    events on its stream (for profiling this shape alone)."""
    t = S.workload_uniform_100k(tile_cells=args.tile_cells, bucket_px=args.bucket_px, form=args.form)
    fr = T.frame(t.params)
    ctx = T.Context(0)
    dX, dY, dZ = ctx.upload(t.X), ctx.upload(t.Y), ctx.upload(t.Z)
    dH = ctx.alloc((t.params.nrows, t.params.ncols), np.float32)
    for _ in range(max(args.warmup, 1)):
        ctx.tile_run_dev(t.params, dX, dY, dZ, t.n, dH)
    ctx.profile(True)
    ctx.profile_read(reset=True)
    s0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.tile_run_dev(t.params, dX, dY, dZ, t.n, dH)
    el = time.perf_counter() - s0
    launches, kern_ms = ctx.profile_read(reset=True)
    ctx.profile(False)
    h = dH.to_host()
    for a in (dX, dY, dZ, dH):
        a.free()
    ctx.close()
    k_ms = kern_ms / max(launches, 1)
    pix = fr["W"] * fr["H"]
    out = {"metric": "DEM cells/s (100k uniform points, 512^2 DEM at 2 m, scale 5, plain form)",
           "value": t.cells * args.steps / el, "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic (seeded uniform LiDAR, LAS int32 at 0.01 m)",
           "config": {"workload": "100k uniform points in 1000 x 1000 m, 512x512 cells at 2 m, scale 5",
                      "points": t.n, "cells": t.cells, "pixels": pix, "form": "plain"},
           "pixels_per_s": pix * args.steps / el, "k_tiles_ms": k_ms}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        ref = O.run(t.params, t.X, t.Y, t.Z, nthreads=min(len(os.sched_getaffinity(0)), 16))
        out["bits_differ_from_oracle"] = int(np.count_nonzero(h.view(np.uint32) != ref["h_fixed"].view(np.uint32)))
    print(json.dumps(compact(out)))


def bench_batched_only(T, S, args):
    """--workload batch: only the batched form over the 256 uneven tiles,
    K graph replays timed (for profiling that form alone)."""
    tiles = S.workload_batch()
    ctx = T.Context(0)
    dev = [(ctx.upload(t.X), ctx.upload(t.Y), ctx.upload(t.Z),
            ctx.alloc((t.params.nrows, t.params.ncols), np.float32)) for t in tiles]
    P = [t.params for t in tiles]
    cols = ([d[0] for d in dev], [d[1] for d in dev], [d[2] for d in dev])
    ns = [t.n for t in tiles]
    hs = [d[3] for d in dev]
    tuned = None
    if not args.no_tune:
        tuned = ctx.autotune(P, *cols, ns, batched=True, reps=2)
    B = ctx.batch_args(P, *cols, ns, hs)  # the C arguments, made once
    for _ in range(max(args.warmup, 1)):
        ctx.batch_run_args(B)
    s0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.batch_run_args(B)
    el = time.perf_counter() - s0
    cells = sum(t.cells for t in tiles)
    pts = sum(ns)
    out = {"metric": "DEM cells/s (256 uneven tiles, batched form)", "value": cells * args.steps / el,
           "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int32",
           "data": "synthetic (seeded uniform / Matern-cluster LiDAR, LAS int32 at 0.01 m)",
           "config": {"workload": "256 uneven tiles, log-normal points (median 200k, 10k-2M), "
                                  "grids 256^2-2048^2 at 0.5 m, one graph-replayed launch",
                      "points": pts, "cells": cells, "form": "batched",
                      "tuned_classes": None if tuned is None else len(tuned)},
           "points_per_s": pts * args.steps / el}
    for d in dev:
        for a in d:
            a.free()
    ctx.close()
    print(json.dumps(out))


def bench_stream(T, S, args):
    """--workload stream: the streamed-terrain pipeline end to end (fBm
    terrain, Hurst 0.8, 0.8 points/m^2 by default, LAS tiles of 1 km with
    5 % overlap duplicates): the driver (LAS reader, one-pass binning with
    border copies, batched engine, GridFloat output) with int64 fixed-point
    and with FP32 two-sum heights; the CPU oracle on about 1 % of the work
    packages is the CPU baseline and the error check.  --stream-shm keeps
    the packages and the output raster in /dev/shm (the LAS files stay on
    disk); each run's raster is sampled and deleted before the next."""
    import shutil
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    driver = os.path.join(ROOT, "terranni_amd", "_lib", "terranni")
    work = tempfile.mkdtemp(prefix="tnni_stream_")
    fast = tempfile.mkdtemp(prefix="tnni_stream_", dir="/dev/shm") if args.stream_shm else work
    ext = args.stream_km * 1000.0
    cs = 0.5
    grid = int(round(ext / cs))
    origin = (500000.0, 4200000.0)
    wp = os.path.join(fast, "wp")
    os.makedirs(wp, exist_ok=True)
    try:
        las = os.path.join(work, "las")
        os.makedirs(las)
        t0 = time.perf_counter()
        files = S.streamed_terrain_las(las, seed=5, extent_m=ext, n_points=args.stream_points, progress=log,
                                       origin=origin)
        n_recs = sum((os.path.getsize(f) - 227) // 20 for f in files)
        gen_s = time.perf_counter() - t0
        log(f"streamed terrain: {len(files)} LAS tiles, {n_recs} records, generated in {gen_s:.1f}s")

        def run(tag, extra):
            out = os.path.join(fast, "out_" + tag)
            os.makedirs(out, exist_ok=True)
            cmd = [driver, f"--tilepath={las}", f"--output={out}", f"--wp-path={wp}",
                   f"--tmp-path={fast}", f"--cell-size={cs}", "--scaling=1", "--site-radius=5",
                   f"--origin-x={origin[0]}", f"--origin-y={origin[1]}", f"--grid-cols={grid}",
                   f"--grid-rows={grid}", f"--wp-width={args.stream_wp_m}",
                   f"--wp-height={args.stream_wp_m}", "--engine=batched", "--2d", "--format=flt",
                   *extra]
            a = time.perf_counter()
            ph, tail = {}, []
            with subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True,
                                  bufsize=1) as pr:
                for line in pr.stderr:
                    line = line.rstrip()
                    tail = (tail + [line])[-20:]
                    log(f"driver {tag}: {line}")
                    if line.startswith("Total ") and ":" in line:
                        k, v = line.split(":", 1)
                        ph[k.replace("Total ", "").replace(" time", "").strip().lower()] = float(v)
            wall = time.perf_counter() - a
            if pr.returncode != 0:
                raise RuntimeError("driver failed: " + "\n".join(tail))
            log(f"driver {tag}: {wall:.2f}s {ph}")
            return wall, ph, out

        # about 1 % of the work packages (at least one), sampled from each run's raster
        import pyoracle as O
        wall64, ph64, out64 = run("fixed64", ["--accum=fixed64"])
        meta = dict(line.split("=", 1) for line in open(os.path.join(wp, "wp.meta")).read().split())
        nc, nr = int(meta["wp_ncols"]), int(meta["wp_nrows"])
        cols, rows = int(meta["cols"]), int(meta["rows"])
        rng = np.random.default_rng(7)
        k = max(1, int(round(0.01 * cols * rows)))
        picks = sorted(int(w) for w in rng.choice(cols * rows, k, replace=False))

        def windows(out):
            m = np.memmap(os.path.join(out, "0.00000.flt"), np.float32, "r", shape=(grid, grid))
            got = {}
            for w in picks:
                r, c = divmod(w, cols)
                j0, j1 = r * nr, min(grid, (r + 1) * nr)
                i0, i1 = c * nc, min(grid, (c + 1) * nc)
                got[w] = np.array(m[grid - j1:grid - j0, i0:i1][::-1])  # rows written top down
            del m
            return got

        s64 = windows(out64)
        shutil.rmtree(out64, ignore_errors=True)
        wall32, ph32, out32 = run("fp32", ["--accum=fp32", "--wp-computed"])
        s32 = windows(out32)
        shutil.rmtree(out32, ignore_errors=True)
        raw = np.dtype([("X", "<i4"), ("Y", "<i4"), ("Z", "<i4"), ("t", "<i4")])
        threads = min(len(os.sched_getaffinity(0)), 16)
        o_s, o_cells, bits, bits32 = 0.0, 0, 0, 0
        e64 = e32 = 0.0
        sq64 = sq32 = 0.0
        nv = 0
        for w in picks:
            r, c = divmod(w, cols)
            pts = np.fromfile(os.path.join(wp, f"r{r}c{c}.raw"), dtype=raw)
            p = T.TileParams(nrows=nr, ncols=nc, cell_size=cs, scale=1, site_radius=5.0,
                             origin_x=origin[0] + c * nc * cs, origin_y=origin[1] + r * nr * cs,
                             las_scale=S.LAS_SCALE,
                             las_offset=(float(meta["offset0"]), float(meta["offset1"]), float(meta["offset2"])))
            a = time.perf_counter()
            ref = O.run(p, pts["X"], pts["Y"], pts["Z"], nthreads=threads)
            o_s += time.perf_counter() - a
            g64, g32 = s64[w], s32[w]
            rf = ref["h_fixed"][:g64.shape[0], :g64.shape[1]]
            r64 = ref["h64"][:g64.shape[0], :g64.shape[1]]
            o_cells += g64.size
            bits += int(np.count_nonzero(g64.view(np.uint32) != rf.view(np.uint32)))
            bits32 += int(np.count_nonzero(g32.view(np.uint32) != g64.view(np.uint32)))
            v = r64 != T.NODATA
            if v.any():
                den = np.maximum(np.abs(r64[v]), 1e-6)
                d64 = np.abs(g64[v] - r64[v]) / den
                d32 = np.abs(g32[v] - r64[v]) / den
                e64, e32 = max(e64, float(d64.max())), max(e32, float(d32.max()))
                sq64 += float(np.sum(d64 * d64))
                sq32 += float(np.sum(d32 * d32))
                nv += int(v.sum())
            log(f"oracle: package r{r}c{c}, {pts.shape[0]} points, {bits} height bits differ so far")
        nni_s = ph64.get("nni", wall64)
        cells = grid * grid
        out = {
            "metric": "DEM cells/s (streamed terrain, driver end to end, batched engine)",
            "value": cells / nni_s, "unit": "cells/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": nni_s * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "int32",
            "data": "synthetic (seeded fBm terrain, Hurst 0.8, LAS int32 at 0.01 m)",
            "config": {"workload": f"streamed terrain {args.stream_km} x {args.stream_km} km, "
                                   f"{grid}^2 DEM at 0.5 m, {len(files)} LAS tiles of 1 km with 5 % "
                                   f"overlap, work packages of {args.stream_wp_m} m",
                       "points": int(n_recs), "cells": cells, "packages": cols * rows,
                       "form": "driver, batched engine"},
            "points_per_s": n_recs / nni_s,
            "end_to_end_cells_per_s": cells / wall64,
            "end_to_end_s": wall64, "generate_s": gen_s,
            "phases_s": {"fixed64": ph64, "fp32": ph32},
            "cpu_baseline": {"value": o_cells / o_s, "unit": "cells/s", "cores": threads,
                             "kind": "port",
                             "sample": f"{k} of {cols * rows} work packages ({o_cells} cells)"},
            "accuracy_sampled": {"fixed64_bits_differ_from_oracle": bits, "fp32_bits_differ_from_fixed64": bits32,
                                 "fixed64_max_rel_vs_fp64": e64,
                                 "fixed64_rms_rel_vs_fp64": float(np.sqrt(sq64 / max(nv, 1))),
                                 "fp32_max_rel_vs_fp64": e32,
                                 "fp32_rms_rel_vs_fp64": float(np.sqrt(sq32 / max(nv, 1)))},
        }
        print(json.dumps(compact(out)))
    finally:
        shutil.rmtree(work, ignore_errors=True)
        if fast != work:
            shutil.rmtree(fast, ignore_errors=True)


def bench_spacetime(T, S, args):
    """--workload spacetime: one work package stepped through time as the
    reference's time loop calls performNNI (nni.cpp:870-944): each call at
    cur_time takes the points of the time-sorted stream within 2r + 1 of it
    (every site of its 2 (int)r + 1 slices), device-resident, as a window of
    one uploaded array.  Delta keeps per point time the 2-D Voronoi layer
    in the space-time cache, combines each slice from the layers within its
    radius and searches only the layer that is new; Delta^2 searches every
    slice.  Both must give the same heights bit for bit; the CPU oracle on
    one call is the CPU baseline and the check."""
    from dataclasses import replace
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    r = args.time_radius
    w = 2.0 * r + 1.0
    first = float(args.st_first)
    curs = [first + k for k in range(args.warmup + args.steps)]
    # survey years so that every call's window is full
    years = np.arange(int(np.floor(first - w)), int(np.ceil(curs[-1] + w)) + 1)
    t0 = time.perf_counter()
    tile = S.make_tile(9, args.st_points_per_year * len(years), args.st_grid, kind="clustered")
    rng = np.random.default_rng(9)
    times = rng.choice(years, tile.n).astype(np.int32)
    o = np.argsort(times, kind="stable")
    X, Y, Z, Tm = (np.ascontiguousarray(a[o]) for a in (tile.X, tile.Y, tile.Z, times))
    p = replace(tile.params)
    cells = p.nrows * p.ncols
    log(f"space-time workload: {tile.n} points over {len(years)} years, {cells} cells, "
        f"generated in {time.perf_counter() - t0:.1f}s")
    ctx = T.Context(0)
    dev = [ctx.upload(a) for a in (X, Y, Z, Tm)]
    dH = ctx.alloc((p.nrows, p.ncols), np.float32)

    def window(cur):
        # integer times: |T - cur| <= w  <=>  ceil(cur - w) <= T <= floor(cur + w)
        a = int(np.searchsorted(Tm, np.int32(np.ceil(cur - w)), side="left"))
        b = int(np.searchsorted(Tm, np.int32(np.floor(cur + w)), side="right"))
        views = [T.DevArray(d.ptr + 4 * a, 4 * (b - a), np.dtype(np.int32), (b - a,), ctx) for d in dev]
        return a, b, views

    windows = {cur: window(cur) for cur in curs}

    def call(cur, cache):
        a, b, (vX, vY, vZ, vT) = windows[cur]
        ctx.tile_run_time_dev(p, vX, vY, vZ, vT, b - a, cur, r, dH, cache=cache)
        return b - a

    res = {}
    heights = {}
    for tag, use in (("delta", True), ("delta_squared", False)):
        cache = ctx.time_cache() if use else None
        for cur in curs[:args.warmup]:
            call(cur, cache)
        lat, pts, hs = [], 0, []
        for cur in curs[args.warmup:]:
            s0 = time.perf_counter()
            pts += call(cur, cache)  # ends in a stream sync
            lat.append(time.perf_counter() - s0)
            hs.append(dH.to_host())
        el = sum(lat)
        res[tag] = {"cells_per_s": cells * args.steps / el, "ms_per_call": el / args.steps * 1e3,
                    "window_points_per_call": pts / args.steps}
        if cache is not None:
            hits, misses = cache.stats()
            res[tag].update({"layers_reused": hits, "layers_searched": misses})
            cache.close()
        heights[tag] = hs
        log(f"space-time {tag}: {res[tag]}")
    same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
               for a, b in zip(heights["delta"], heights["delta_squared"]))
    out = {
        "metric": f"DEM cells/s (space-time tile {p.ncols}^2, time radius {r:g}, Delta cache)",
        "value": res["delta"]["cells_per_s"], "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["delta"]["ms_per_call"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": f"synthetic (seeded Matern-cluster LiDAR, survey years {years[0]}-{years[-1]}, "
                "LAS int32 at 0.01 m)",
        "config": {"workload": f"space-time tile {p.ncols}x{p.nrows} cells at {p.cell_size} m, "
                               f"{tile.n} points over {len(years)} years, time radius {r:g}, "
                               f"cur_time {curs[args.warmup]:g}..{curs[-1]:g}",
                   "points": tile.n, "cells": cells, "form": "plain, device-resident windows"},
        "delta": res["delta"], "delta_squared": res["delta_squared"],
        "delta_vs_delta_squared": res["delta"]["cells_per_s"] / res["delta_squared"]["cells_per_s"],
        "delta_equals_delta_squared": same,
    }
    if not args.no_cpu:
        cur = curs[-1]
        a, b, _ = windows[cur]
        threads = min(len(os.sched_getaffinity(0)), 16)
        s0 = time.perf_counter()
        ref = O.run_time(p, X[a:b], Y[a:b], Z[a:b], Tm[a:b], cur, r, nthreads=threads)
        o_s = time.perf_counter() - s0
        out["cpu_baseline"] = {"value": cells / o_s, "unit": "cells/s", "cores": threads, "kind": "port",
                               "sample": f"one call (cur_time {cur:g}, all slices) of the same tile"}
        out["matches_oracle"] = bool(np.array_equal(heights["delta"][-1].view(np.uint32),
                                                    ref["h_fixed"].view(np.uint32)))
    for d in dev + [dH]:
        d.free()
    ctx.close()
    print(json.dumps(out))


def load_traffic(kernel="k_tiles<0, 0, false, false, false>"):
    """HBM bytes per k_tiles launch on this workload, from the committed
    FETCH_SIZE / WRITE_SIZE summary (profiles/traffic_single_4m.json)."""
    path = os.path.join(ROOT, "profiles", "traffic_single_4m.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k["bytes"]


VALU_PEAK = 256 * 2.4e9  # wave64 VALU instructions per second: 1 per CU per cycle


def load_valu(kernel="k_tiles<0, 0, false, false, false>"):
    """VALU wave-instructions per k_tiles launch on this workload, from the
    committed SQ_INSTS_VALU summary (profiles/sq_counters_single_4m.json)."""
    path = os.path.join(ROOT, "profiles", "sq_counters_single_4m.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k.get("SQ_INSTS_VALU")


def load_lds(kernel="k_tiles<0, 0, false, false, false>"):
    """LDS activity of a k_tiles launch on this workload, from the committed
    LDS counter summary (profiles/lds_counters_single_4m.json): the share of
    CU-cycles the LDS is busy (peak 128 B per CU per cycle) and the share of
    those stalled by bank conflicts."""
    path = os.path.join(ROOT, "profiles", "lds_counters_single_4m.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f).get("kernels", {}).get(kernel)
    if k is None:
        return None
    return {"active_frac": k["lds_active_frac"], "bank_conflict_frac": k["lds_bank_conflict_frac"],
            "peak_bytes_per_cu_clk": 128, "out_of_range_accesses": k["SQ_LDS_MEM_VIOLATIONS"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=4096)
    ap.add_argument("--points", type=int, default=4_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-cells", type=int, default=0, help="cell-tile side (0 = auto)")
    ap.add_argument("--bucket-px", type=int, default=0, help="bucket side in pixels (0 = auto)")
    ap.add_argument("--form", type=int, default=0,
                    help="0: automatic, 1: fused, 2: split (label raster, then a gather pass), "
                         "3: scatter (label raster, then each pixel spread over its cells)")
    ap.add_argument("--workload", choices=("single", "batch", "stream", "spacetime", "ref100k"),
                    default="single",
                    help="single: the 4096^2 tile (default line, plus the batched and served "
                         "comparisons); batch: only the batched 256-tile form; stream: the "
                         "streamed-terrain driver pipeline (down-scaled); spacetime: one work "
                         "package stepped through time, Delta cache against Delta^2; ref100k: "
                         "the reference's CPU case (100k uniform points, 512^2 at 2 m, scale 5) alone")
    ap.add_argument("--time-radius", type=float, default=2.0, help="spacetime: time radius")
    ap.add_argument("--st-grid", type=int, default=1024, help="spacetime: tile side in cells")
    ap.add_argument("--st-points-per-year", type=int, default=400_000)
    ap.add_argument("--st-first", type=float, default=2010.0, help="spacetime: first cur_time")
    ap.add_argument("--stream-km", type=float, default=4.0, help="streamed terrain side (km)")
    ap.add_argument("--stream-points", type=int, default=12_000_000)
    ap.add_argument("--stream-wp-m", type=int, default=1000, help="work package side (m)")
    ap.add_argument("--stream-shm", action="store_true",
                    help="stream: packages and output raster in /dev/shm (the LAS files on disk)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (built-in bucket and cell-tile sizes)")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-ref-case", action="store_true",
                    help="skip the 100k-point 512^2 case (profiler runs of the 4M tile)")
    ap.add_argument("--served-rates", default="50,100,200,500,1000,2000,0",
                    help="offered tiles/s, HBM-resident requests (0: saturated, back-to-back "
                         "submits)")
    ap.add_argument("--served-host-rates", default="200,1000,0",
                    help="offered tiles/s, host-buffer requests")
    ap.add_argument("--served-requests", type=int, default=2000,
                    help="requests per arrival rate")
    ap.add_argument("--record-dir", default=os.path.join(ROOT, "bench_records"),
                    help="where the whole unrounded record (bench_full.json, with the autotune "
                         "tables) is written")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters against known traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); tiles are "
              "independent, so more GPUs would each run their own share", file=sys.stderr)
        sys.exit(2)

    import terranni_amd as T
    import terranni_amd.synth as S

    if args.workload == "batch":
        bench_batched_only(T, S, args)
        return
    if args.workload == "stream":
        bench_stream(T, S, args)
        return
    if args.workload == "spacetime":
        bench_spacetime(T, S, args)
        return
    if args.workload == "ref100k":
        bench_ref100k(T, S, args)
        return
    t0 = time.perf_counter()
    tile = S.workload_single_4m(seed=1, n=args.points, grid=args.grid,
                                tile_cells=args.tile_cells, bucket_px=args.bucket_px, form=args.form)
    log(f"generated {tile.n} points, {tile.cells} cells in {time.perf_counter() - t0:.1f}s")
    fr = T.frame(tile.params)

    ctx = T.Context(0)
    dX, dY, dZ = ctx.upload(tile.X), ctx.upload(tile.Y), ctx.upload(tile.Z)
    dH = ctx.alloc((tile.params.nrows, tile.params.ncols), np.float32)

    for _ in range(args.warmup):
        ctx.tile_run_dev(tile.params, dX, dY, dZ, tile.n, dH)
    ctx.profile(True)
    ctx.profile_read(reset=True)
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        ctx.tile_run_dev(tile.params, dX, dY, dZ, tile.n, dH)  # ends in a stream sync
        lat.append(time.perf_counter() - s0)
    elapsed = time.perf_counter() - t_start
    launches, kern_ms = ctx.profile_read(reset=True)
    ctx.profile(False)
    h = dH.to_host()
    # the same tile with sizes the autotuner measured for it (reported
    # beside `value`, which uses the built-in sizes the committed profiles
    # were taken with)
    tuned = None
    if not args.no_tune:
        tuning = ctx.autotune([tile.params], [dX], [dY], [dZ], [tile.n], batched=False, reps=3)
        log(f"autotuned single tile: {tuning}")
        for _ in range(args.warmup):
            ctx.tile_run_dev(tile.params, dX, dY, dZ, tile.n, dH)
        s0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.tile_run_dev(tile.params, dX, dY, dZ, tile.n, dH)
        el_t = time.perf_counter() - s0
        same_t = bool(np.array_equal(dH.to_host().view(np.uint32), h.view(np.uint32)))
        tuned = {k: tuning[0][k] for k in ("bucket_px", "tile_cells", "wg_per_cu")}
        tuned.update({"cells_per_s": tile.cells * args.steps / el_t,
                      "ms_per_step": el_t / args.steps * 1e3, "same_heights": same_t})
    # PCIe-inclusive: host buffers in, heights back to host memory
    t_h = time.perf_counter()
    for _ in range(3):
        ctx.tile_run(tile.params, tile.X, tile.Y, tile.Z)
    pcie_cells_s = 3 * tile.cells / (time.perf_counter() - t_h)
    # (not in a counter calibration run, whose one probe copy must be the
    # only large copy)
    copy_gbs = copy_bandwidth(ctx) if not args.copy_probe_mb else None
    if args.copy_probe_mb:
        nb = args.copy_probe_mb << 20
        a, c = ctx.alloc((nb,), np.uint8), ctx.alloc((nb,), np.uint8)
        ctx.copy(c, a)
        a.free()
        c.free()

    cells_s = tile.cells * args.steps / elapsed
    pts_s = tile.n * args.steps / elapsed
    k_ms = kern_ms / max(launches, 1)
    alg_bytes = 12 * tile.n + 4 * tile.cells  # points in (X, Y, Z int32) + FP32 heights out
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9

    out = {
        "metric": "DEM cells/s (single 4096^2 tile, 4M clustered points, plain form)",
        "value": cells_s,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded Matern-cluster LiDAR, LAS int32 at 0.01 m)",
        "config": {"workload": "single tile 4096x4096 cells at 0.5 m, scale 1, 4M clustered points",
                   "points": tile.n, "cells": tile.cells, "pixels": fr["W"] * fr["H"],
                   "form": "plain", "sizes": "built-in (bucket and cell-tile sides)"},
        "tuned": tuned,
        "points_per_s": pts_s,
        "pcie_inclusive_cells_per_s": pcie_cells_s,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(),
                     "peak_measured_copy": copy_gbs,
                     "kernel": "k_tiles", "kernel_ms": k_ms, "alg_bytes": alg_bytes},
    }
    valu = load_valu()
    if valu:
        # the bound that matters for this kernel: integer VALU issue
        rate = valu / (k_ms * 1e-3)
        out["roofline"]["valu"] = {"instructions": valu, "achieved_per_s": rate,
                                   "peak_per_s": VALU_PEAK, "frac": rate / VALU_PEAK}
    lds = load_lds()
    if lds:
        out["roofline"]["lds"] = lds
    for d in (dX, dY, dZ, dH):
        d.free()
    pool, table, choices = None, None, None
    if not args.no_ref_case:
        out["uniform_100k"] = bench_uniform_100k(T, S, ctx, cpu=not args.no_cpu)
    if not args.no_batch:
        out["batched_vs_plain"], pool, table, choices = bench_batch(T, S, ctx, tune=not args.no_tune,
                                                                     cpu=not args.no_cpu)
    if not args.no_served:
        if pool is None:
            pool = S.workload_batch()
        rates = [float(r) for r in args.served_rates.split(",") if r]
        host_rates = [float(r) for r in args.served_host_rates.split(",") if r]
        out["served"] = {
            "workload": "Poisson arrivals of tiles drawn from the 256 uneven tiles; rate 0: "
                        "saturated (back-to-back submits); latency submit -> results",
            "hbm_resident": bench_served(T, pool, rates, args.served_requests, True,
                                         tune_table=table),
            "host_buffers": bench_served(T, pool, host_rates, args.served_requests, False,
                                         tune_table=table),
        }
    if not args.no_cpu:
        c_cells, threads, per, reps, ref = cpu_baseline(tile)
        out["cpu_baseline"] = {"value": c_cells, "unit": "cells/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the same tile, full, {reps} run(s), {per:.2f} s each",
                               "single_thread_value": cpu_single_thread(tile),
                               "cpu_model": cpu_model()}
        valid = ref["h64"] != T.NODATA
        # relative to the FP64 mean, floored at 1 micrometre: an exact integer
        # sum of 0 (cancellation) leaves FP64 rounding noise in the oracle mean
        rel = np.abs(h[valid] - ref["h64"][valid]) / np.maximum(np.abs(ref["h64"][valid]), 1e-6)
        out["accuracy"] = {"max_rel_err": float(rel.max()),
                           "rms_rel_err": float(np.sqrt(np.mean(rel ** 2))),
                           "bits_differ_from_oracle": int(np.count_nonzero(
                               h.view(np.uint32) != ref["h_fixed"].view(np.uint32))),
                           "nodata_match": bool(np.array_equal(valid, h != T.NODATA)),
                           "stolen_pairs": ref["S"], "candidate_pairs": ref["K"],
                           "stolen_pairs_per_s": ref["S"] / (elapsed / args.steps)}
    ctx.close()
    # the secondary figures once more, short, as the line's last key: a log
    # that keeps only the line's last characters still holds them
    summ = {"k_tiles_ms": out["roofline"]["kernel_ms"]}
    if "valu" in out["roofline"]:
        summ["k_tiles_valu_frac"] = out["roofline"]["valu"]["frac"]
    if "uniform_100k" in out:
        summ["ref100k_ms"] = out["uniform_100k"]["gpu_ms"]
    bvp = out.get("batched_vs_plain")
    if bvp:
        summ["batched_speedup"] = bvp["speedup"]
        summ["batched_cells_per_s"] = bvp["batched_cells_per_s"]
        summ["plain_cells_per_s"] = bvp["plain_cells_per_s"]
    srv = out.get("served")
    if srv:
        for kind in ("hbm_resident", "host_buffers"):
            summ[f"served_{kind}_p50_p99_ms"] = {
                ("saturated" if r["rate"] == 0 else str(int(r["rate"]))): [r["p50_ms"], r["p99_ms"]]
                for r in srv.get(kind, [])}
    out["summary"] = summ
    full = dict(out)
    if choices is not None:
        full["autotune_choices"] = choices
    write_full_record(full, "bench_full.json", args.record_dir)
    print(json.dumps(compact(out)))


if __name__ == "__main__":
    main()
