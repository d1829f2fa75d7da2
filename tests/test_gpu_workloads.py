"""The batched, served and streamed workloads at their stated sizes, on the
GPU, against the CPU oracle.

- 256 uneven tiles (log-normal point counts, median 200k, 10k-2M; grids of
  256^2-2048^2 at 0.5 m; half uniform, half clustered) in one batched
  launch: every tile's labels and squared distances bit-exact, heights
  bit-equal to the oracle's exact formula and within 1e-5 of its FP64 mean.
  The plain one-tile-per-launch form gives the same bits.
- A served stream: Poisson arrivals of tiles drawn from that pool, on device
  buffers: every result equals the batched and plain ones, and requests
  overlap on the device.
- Streamed terrain, down-scaled: fBm terrain (Hurst 0.8) over 2 x 2 km, LAS
  tiles of 1 km with 5 % overlap, through the driver (binning with border
  copies, batched engine, GridFloat out) with int64 fixed-point and FP32
  two-sum heights; sampled work packages against the oracle.
(terranni.cpp:394-426 is the tile loop these forms replace.)"""
import os
import shutil
import subprocess
import tempfile
import time

import numpy as np
import pytest

import pyoracle as O
import terranni_amd as T
import terranni_amd.synth as S
from helpers import REL_TOL, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
THREADS = min(len(os.sched_getaffinity(0)), 16)


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pool(ctx):
    _say("generating the 256 uneven tiles")
    t0 = time.perf_counter()
    tiles = S.workload_batch()
    dev = [(ctx.upload(t.X), ctx.upload(t.Y), ctx.upload(t.Z)) for t in tiles]
    _say(f"  {len(tiles)} tiles, {sum(t.n for t in tiles)} points, {time.perf_counter() - t0:.1f}s")
    yield tiles, dev
    for row in dev:
        for a in row:
            a.free()


@pytest.fixture(scope="module")
def batched(ctx, pool):
    """Every tile's heights, labels and squared distances from one batched
    launch."""
    tiles, dev = pool
    outs = []
    for t in tiles:
        f = T.frame(t.params)
        outs.append((ctx.alloc((t.params.nrows, t.params.ncols), np.float32),
                     ctx.alloc((f["H"], f["W"]), np.int32), ctx.alloc((f["H"], f["W"]), np.uint32)))
    _say("batched launch of the 256 tiles")
    ctx.batch_run_dev([t.params for t in tiles], [d[0] for d in dev], [d[1] for d in dev],
                      [d[2] for d in dev], [t.n for t in tiles], [o[0] for o in outs],
                      [o[1] for o in outs], [o[2] for o in outs])
    host = []
    for o in outs:
        host.append(tuple(a.to_host() for a in o))
        for a in o:
            a.free()
    return host


@pytest.mark.timeout(600)
def test_batched_256_uneven_tiles_match_the_oracle(pool, batched):
    tiles, _ = pool
    emax, sq, ncells, nlab = 0.0, 0.0, 0, 0
    for k, (t, (h, lab, d2)) in enumerate(zip(tiles, batched)):
        ref = O.run(t.params, t.X, t.Y, t.Z, nthreads=THREADS)
        bad = int((lab != ref["labels"]).sum())
        assert bad == 0, f"tile {k} ({t.n} points, {t.params.nrows}^2, {t.kind}): {bad} labels differ"
        assert np.array_equal(d2, ref["d2"]), f"tile {k}: squared distances differ"
        assert np.array_equal(h.view(np.uint32), ref["h_fixed"].view(np.uint32)), f"tile {k}: heights differ"
        e, rms, same = rel_err(h, ref["h64"])
        assert same, f"tile {k}: NODATA cells differ"
        assert e <= REL_TOL, (k, e)
        emax = max(emax, e)
        sq += rms * rms * h.size
        ncells += h.size
        nlab += lab.size
    _say(f"256 tiles: {nlab} labels and {ncells} heights checked; height error vs FP64: "
         f"max {emax:.2e}, rms {np.sqrt(sq / ncells):.2e}")


@pytest.mark.timeout(300)
def test_plain_form_equals_batched_on_the_256_tiles(ctx, pool, batched):
    tiles, dev = pool
    side = max(max(T.frame(t.params)["W"], T.frame(t.params)["H"]) for t in tiles)
    cells = max(t.params.nrows * t.params.ncols for t in tiles)
    dh = ctx.alloc((cells,), np.float32)
    dl = ctx.alloc((side * side,), np.int32)
    dd = ctx.alloc((side * side,), np.uint32)
    for k, (t, d, (h, lab, d2)) in enumerate(zip(tiles, dev, batched)):
        ctx.tile_run_dev(t.params, d[0], d[1], d[2], t.n, dh, dl, dd)
        hp = dh.to_host()[:h.size].reshape(h.shape)
        assert np.array_equal(hp.view(np.uint32), h.view(np.uint32)), f"tile {k}: heights differ"
        assert np.array_equal(dl.to_host()[:lab.size].reshape(lab.shape), lab), f"tile {k}: labels differ"
        assert np.array_equal(dd.to_host()[:d2.size].reshape(d2.shape), d2), f"tile {k}: d2 differ"
    for a in (dh, dl, dd):
        a.free()


@pytest.mark.timeout(300)
def test_served_poisson_stream_from_the_pool(ctx, pool, batched):
    """400 Poisson arrivals at 1000 tiles/s, each a tile drawn from the
    pool, submitted on device buffers as they arrive: each result equals
    the batched one (and so the plain one and the oracle), and tiles share
    the device (continuous batching)."""
    tiles, dev = pool
    rng = np.random.default_rng(11)
    n_req, rate = 400, 1000.0
    pick = rng.integers(0, len(tiles), n_req)
    due = np.cumsum(rng.exponential(1.0 / rate, n_req))
    outs, pend = [], []
    for k in pick:
        t = tiles[k]
        f = T.frame(t.params)
        outs.append((ctx.alloc((t.params.nrows, t.params.ncols), np.float32),
                     ctx.alloc((f["H"], f["W"]), np.int32), ctx.alloc((f["H"], f["W"]), np.uint32)))
    ctx.serve_stats_reset()
    _say(f"served: {n_req} arrivals at {rate:.0f} tiles/s")
    t0 = time.perf_counter()
    for i, k in enumerate(pick):
        while time.perf_counter() - t0 < due[i]:
            pass
        d, o = dev[k], outs[i]
        pend.append(ctx.submit_dev(tiles[k].params, d[0], d[1], d[2], tiles[k].n, o[0], o[1], o[2]))
    lat = []
    for p in pend:
        p.wait()
        lat.append(p.latency_s)
    stats = ctx.serve_stats()
    _say(f"  done in {time.perf_counter() - t0:.2f}s, p50 {np.percentile(lat, 50) * 1e3:.2f} ms, "
         f"p99 {np.percentile(lat, 99) * 1e3:.2f} ms, most tiles in flight {stats['max_inflight_tiles']}, "
         f"most batches in flight {stats['max_inflight_batches']}")
    for i, k in enumerate(pick):
        h, lab, d2 = batched[k]
        o = outs[i]
        assert np.array_equal(o[0].to_host().view(np.uint32), h.view(np.uint32)), f"request {i}: heights"
        assert np.array_equal(o[1].to_host(), lab), f"request {i}: labels"
        assert np.array_equal(o[2].to_host(), d2), f"request {i}: squared distances"
        for a in o:
            a.free()
    assert stats["max_inflight_tiles"] >= 2


@pytest.mark.timeout(600)
def test_streamed_terrain_driver_sampled_packages_match_the_oracle():
    """fBm terrain over 2 x 2 km (3.2M points, the full workload's density
    of 0.8 points/m^2), 4000^2 DEM at 0.5 m, work packages of 500 m: the
    driver with int64 fixed-point and with FP32 two-sum heights; three
    sampled packages bit-equal to the oracle's exact formula (int64) and
    within 1e-5 of its FP64 mean (both)."""
    driver = os.path.join(ROOT, "terranni_amd", "_lib", "terranni")
    work = tempfile.mkdtemp(prefix="tnni_terrain_")
    ext, cs, wp_m = 2000.0, 0.5, 500
    grid = int(round(ext / cs))
    origin = (500000.0, 4200000.0)
    try:
        las = os.path.join(work, "las")
        os.makedirs(las)
        _say("terrain: generating LAS tiles")
        S.streamed_terrain_las(las, seed=5, extent_m=ext, n_points=3_200_000, origin=origin)

        def run(tag, extra):
            out = os.path.join(work, "out_" + tag)
            os.makedirs(out, exist_ok=True)
            cmd = [driver, f"--tilepath={las}", f"--output={out}", f"--wp-path={work}/wp",
                   f"--tmp-path={work}", f"--cell-size={cs}", "--scaling=1", "--site-radius=5",
                   f"--origin-x={origin[0]}", f"--origin-y={origin[1]}", f"--grid-cols={grid}",
                   f"--grid-rows={grid}", f"--wp-width={wp_m}", f"--wp-height={wp_m}",
                   "--engine=batched", "--2d", "--format=flt", *extra]
            _say("terrain: driver", tag)
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
            assert r.returncode == 0, r.stderr[-3000:]
            return np.fromfile(os.path.join(out, "0.00000.flt"), np.float32).reshape(grid, grid)[::-1]

        os.makedirs(os.path.join(work, "wp"))
        h64 = run("fixed64", ["--accum=fixed64"])
        h32 = run("fp32", ["--accum=fp32", "--wp-computed"])
        meta = dict(line.split("=", 1) for line in open(os.path.join(work, "wp", "wp.meta")).read().split())
        nc, nr = int(meta["wp_ncols"]), int(meta["wp_nrows"])
        cols, rows = int(meta["cols"]), int(meta["rows"])
        raw = np.dtype([("X", "<i4"), ("Y", "<i4"), ("Z", "<i4"), ("t", "<i4")])
        rng = np.random.default_rng(7)
        e64 = e32 = 0.0
        for w in rng.choice(cols * rows, 3, replace=False):
            r, c = divmod(int(w), cols)
            pts = np.fromfile(os.path.join(work, "wp", f"r{r}c{c}.raw"), dtype=raw)
            p = T.TileParams(nrows=nr, ncols=nc, cell_size=cs, scale=1, site_radius=5.0,
                             origin_x=origin[0] + c * nc * cs, origin_y=origin[1] + r * nr * cs,
                             las_scale=S.LAS_SCALE,
                             las_offset=(float(meta["offset0"]), float(meta["offset1"]),
                                         float(meta["offset2"])))
            ref = O.run(p, pts["X"], pts["Y"], pts["Z"], nthreads=THREADS)
            j1, i1 = min(grid, (r + 1) * nr), min(grid, (c + 1) * nc)
            g64 = h64[r * nr:j1, c * nc:i1]
            g32 = h32[r * nr:j1, c * nc:i1]
            rf = ref["h_fixed"][:j1 - r * nr, :i1 - c * nc]
            r64 = ref["h64"][:j1 - r * nr, :i1 - c * nc]
            _say(f"terrain: package r{r}c{c}, {pts.shape[0]} points")
            assert np.array_equal(g64.view(np.uint32), rf.view(np.uint32)), f"package r{r}c{c}"
            for g in (g64, g32):
                e, _, same = rel_err(np.ascontiguousarray(g), r64)
                assert same and e <= REL_TOL, (r, c, e)
            e64 = max(e64, rel_err(np.ascontiguousarray(g64), r64)[0])
            e32 = max(e32, rel_err(np.ascontiguousarray(g32), r64)[0])
        valid = h64 != T.NODATA
        assert np.array_equal(valid, h32 != T.NODATA) and valid.mean() > 0.99
        _say(f"terrain: max rel error vs FP64, int64 {e64:.2e}, FP32 two-sum {e32:.2e}; "
             f"FP32 cells equal to int64: {np.mean(h32[valid] == h64[valid]):.6f}")
    finally:
        shutil.rmtree(work, ignore_errors=True)
