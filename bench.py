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


def bench_batch(T, S, ctx, reps: int = 3, tune: bool = True):
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

    def plain():
        for t, d in zip(tiles[:8], dev[:8]):
            ctx.tile_run_dev(t.params, d[0], d[1], d[2], t.n, d[3])
        lat = []
        s0 = time.perf_counter()
        for t, d in zip(tiles, dev):
            a = time.perf_counter()
            ctx.tile_run_dev(t.params, d[0], d[1], d[2], t.n, d[3])
            lat.append(time.perf_counter() - a)
        return time.perf_counter() - s0, lat, [h.to_host() for h in hs]

    def batched():
        ctx.batch_run_dev(P, *cols, ns, hs)
        s0 = time.perf_counter()
        for _ in range(reps):
            ctx.batch_run_dev(P, *cols, ns, hs)
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

    def pct(v):
        return {"p50": float(np.percentile(v, 50) * 1e3), "p99": float(np.percentile(v, 99) * 1e3)}

    return {
        "workload": "256 uneven tiles, log-normal points (median 200k, 10k-2M), grids "
                    "256^2-2048^2 at 0.5 m, half uniform / half clustered",
        "tiles": len(tiles), "points": pts, "cells": cells,
        "plain_cells_per_s": cells / plain_s, "plain_points_per_s": pts / plain_s,
        "plain_tile_latency_ms": pct(lat),
        "batched_cells_per_s": cells / batch_s, "batched_points_per_s": pts / batch_s,
        "batched_ms": batch_s * 1e3,
        "speedup": plain_s / batch_s,
        "tuned": {"classes_batched": len(tuned_b), "classes_plain": len(tuned_p),
                  "tuning_s": tune_s,
                  "plain_cells_per_s": cells / plain_t_s, "plain_tile_latency_ms": pct(lat_t),
                  "batched_cells_per_s": cells / batch_t_s, "batched_ms": batch_t_s * 1e3,
                  "speedup": plain_t_s / batch_t_s,
                  "batched_choices": [{k: e[k] for k in ("bucket_px", "tile_cells", "wg_per_cu",
                                                          "ms_default", "ms_tuned")}
                                      for e in tuned_b]} if tune else None,
        "batched_equals_plain": bool(same(plain_out, batch_out) and same(plain_out, plain_t_out)
                                     and same(plain_out, batch_t_out)),
    }, tiles, ctx.tune_table()


def bench_served(T, tiles, rates, per_rate, resident: bool, seed=4, tune_table=None):
    """Served form: Poisson arrivals at each rate of tiles drawn from the
    uneven-tile pool; one thread submits, another waits in order.  resident:
    the tiles' points and the output rasters live in HBM (tnni_submit_dev);
    otherwise host buffers in and out (tnni_submit).  Latency is submit ->
    results in device memory (resident) or in the caller's host buffers,
    timed by the engine."""
    import threading
    out = []
    rng = np.random.default_rng(seed)
    ring = 192  # output rasters in flight (resident)
    with T.Context(0, n_slots=3, max_batch_tiles=64) as c:
        if tune_table:
            c.tune_load(tune_table)
        if resident:
            dev = [(c.upload(t.X), c.upload(t.Y), c.upload(t.Z)) for t in tiles]
            big = max(t.cells for t in tiles)
            outs = [c.alloc((big,), np.float32) for _ in range(ring)]

        def submit(k, i):
            t = tiles[i]
            if resident:
                d = dev[i]
                return c.submit_dev(t.params, d[0], d[1], d[2], t.n, outs[k % ring])
            return c.submit(t.params, t.X, t.Y, t.Z)

        for p in [submit(k, k) for k in range(16)]:  # warm the slots and graphs
            p.wait()
        for rate in rates:
            picks = rng.integers(0, len(tiles), per_rate)
            gaps = rng.exponential(1.0 / rate, per_rate)
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
            out.append({"rate_tiles_per_s": rate, "requests": per_rate,
                        "achieved_tiles_per_s": per_rate / span, "cells_per_s": cells / span,
                        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                                       "p99": float(np.percentile(lat, 99) * 1e3)},
                        "max_inflight_tiles": st["max_inflight_tiles"],
                        "max_inflight_batches": st["max_inflight_batches"]})
            log(f"served ({'HBM-resident' if resident else 'host buffers'}) rate {rate}/s: {out[-1]}")
        if resident:
            for d in dev:
                for a in d:
                    a.free()
            for o in outs:
                o.free()
    return out


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
    for _ in range(max(args.warmup, 1)):
        ctx.batch_run_dev(P, *cols, ns, hs)
    s0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.batch_run_dev(P, *cols, ns, hs)
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


def load_traffic(kernel="k_tiles<0, 0>"):
    """HBM bytes per k_tiles launch on this workload, from the committed
    FETCH_SIZE / WRITE_SIZE summary (profiles/traffic_single_4m.json)."""
    path = os.path.join(ROOT, "profiles", "traffic_single_4m.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k["bytes"]


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
    ap.add_argument("--workload", choices=("single", "batch"), default="single",
                    help="single: the 4096^2 tile (default line, plus the batched and served "
                         "comparisons); batch: only the batched 256-tile form")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (built-in bucket and cell-tile sizes)")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-rates", default="50,200,500,1000,2000")
    ap.add_argument("--served-requests", type=int, default=300,
                    help="requests per arrival rate")
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
    t0 = time.perf_counter()
    tile = S.workload_single_4m(seed=1, n=args.points, grid=args.grid,
                                tile_cells=args.tile_cells, bucket_px=args.bucket_px)
    log(f"generated {tile.n} points, {tile.cells} cells in {time.perf_counter() - t0:.1f}s")
    fr = T.frame(tile.params)

    ctx = T.Context(0)
    dX, dY, dZ = ctx.upload(tile.X), ctx.upload(tile.Y), ctx.upload(tile.Z)
    dH = ctx.alloc((tile.params.nrows, tile.params.ncols), np.float32)

    tuning = None
    if not args.no_tune:
        tuning = ctx.autotune([tile.params], [dX], [dY], [dZ], [tile.n], batched=False, reps=3)
        log(f"autotuned single tile: {tuning}")
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
    # PCIe-inclusive: host buffers in, heights back to host memory
    t_h = time.perf_counter()
    for _ in range(3):
        ctx.tile_run(tile.params, tile.X, tile.Y, tile.Z)
    pcie_cells_s = 3 * tile.cells / (time.perf_counter() - t_h)
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
                   "form": "plain",
                   "tuning": (None if tuning is None else
                              {k: tuning[0][k] for k in ("bucket_px", "tile_cells", "wg_per_cu",
                                                         "ms_default", "ms_tuned")})},
        "points_per_s": pts_s,
        "pcie_inclusive_cells_per_s": pcie_cells_s,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(),
                     "kernel": "k_tiles", "kernel_ms": k_ms, "alg_bytes": alg_bytes},
    }
    for d in (dX, dY, dZ, dH):
        d.free()
    pool, table = None, None
    if not args.no_batch:
        out["batched_vs_plain"], pool, table = bench_batch(T, S, ctx, tune=not args.no_tune)
    if not args.no_served:
        if pool is None:
            pool = S.workload_batch()
        rates = [float(r) for r in args.served_rates.split(",") if r]
        out["served"] = {
            "workload": "Poisson arrivals of tiles drawn from the 256 uneven tiles",
            "hbm_resident": bench_served(T, pool, rates, args.served_requests, True,
                                         tune_table=table),
            "host_buffers": bench_served(T, pool, rates[:2], args.served_requests, False,
                                         tune_table=table),
        }
    if not args.no_cpu:
        c_cells, threads, per, reps, ref = cpu_baseline(tile)
        out["cpu_baseline"] = {"value": c_cells, "unit": "cells/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the same tile, full, {reps} run(s), {per:.2f} s each"}
        valid = ref["h64"] != T.NODATA
        # relative to the FP64 mean, floored at 1 micrometre: an exact integer
        # sum of 0 (cancellation) leaves FP64 rounding noise in the oracle mean
        rel = np.abs(h[valid] - ref["h64"][valid]) / np.maximum(np.abs(ref["h64"][valid]), 1e-6)
        out["accuracy"] = {"max_rel_err": float(rel.max()),
                           "rms_rel_err": float(np.sqrt(np.mean(rel ** 2))),
                           "bits_differ_from_oracle": int(np.count_nonzero(
                               h.view(np.uint32) != ref["h_fixed"].view(np.uint32))),
                           "nodata_match": bool(np.array_equal(valid, h != T.NODATA)),
                           "stolen_pairs": ref["S"], "candidate_pairs": ref["K"]}
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
