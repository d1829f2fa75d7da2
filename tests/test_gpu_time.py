"""Space-time tiles on the GPU (performNNI with timeRadius > 0,
nni.cpp:902-936): every slice's Voronoi with the sites' time distances, the
queries' time offsets, and the sums over all slices, bit-equal to the
oracle's restatement (heights: int64 sums and the S0 integer path exactly,
FP32 two-sum within 1e-5 of the FP64 mean)."""
import dataclasses

import numpy as np
import pytest

import pyoracle as O
import terranni_amd as T
import terranni_amd.synth as S
from helpers import REL_FLOOR, REL_TOL

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


CASES = [
    # seed, points, grid, kind, scale, algorithm, accumulation, radius, time
    (1, 20000, 96, "clustered", 1, T.ALG_PLANES, T.ACC_FIXED64, 1.0, 2002.0),
    (2, 15000, 64, "uniform", 1, T.ALG_PLANES, T.ACC_FIXED64, 2.0, 2001.0),
    (3, 15000, 64, "uniform", 1, T.ALG_CONES, T.ACC_FIXED64, 1.5, 2002.0),
    (4, 8000, 48, "clustered", 2, T.ALG_PLANES, T.ACC_S0, 1.0, 2002.0),
    (5, 8000, 48, "uniform", 1, T.ALG_CONES, T.ACC_FP32_2SUM, 2.0, 2002.5),
    (6, 5000, 40, "uniform", 1, T.ALG_PLANES, T.ACC_FIXED64, 0.0, 2002.0),
    (7, 12000, 64, "clustered", 1, T.ALG_PLANES, T.ACC_FIXED64, 3.0, 2003.0),
]


@pytest.mark.parametrize("order", ["unsorted", "sorted"])
@pytest.mark.parametrize("seed,n,grid,kind,scale,alg,mode,r,cur", CASES)
def test_space_time_tile_matches_oracle(ctx, seed, n, grid, kind, scale, alg, mode, r, cur, order):
    """unsorted: points in random time order, every slice searched; sorted:
    points sorted by time, searched once per time as 2-D layers and every
    slice combined from them."""
    _say(f"space-time seed={seed} n={n} grid={grid} alg={alg} mode={mode} r={r} {order}")
    t = S.make_tile(seed, n, grid, scale=scale, kind=kind, algorithm=alg, accum_mode=mode,
                    point_min=-60, multiplier=1000)
    times = np.random.default_rng(seed).integers(1998, 2007, t.n).astype(np.int32)
    X, Y, Z = t.X, t.Y, t.Z
    if order == "sorted":
        o = np.argsort(times, kind="stable")
        X, Y, Z, times = X[o], Y[o], Z[o], times[o]
    h = ctx.tile_run_time(t.params, X, Y, Z, times, cur, r)
    ref = O.run_time(t.params, X, Y, Z, times, cur, r, s0=mode == T.ACC_S0)
    if mode == T.ACC_FIXED64:
        assert np.array_equal(h.view(np.uint32), ref["h_fixed"].view(np.uint32))
    elif mode == T.ACC_S0:
        assert np.array_equal(h.view(np.uint32), ref["h_s0"].view(np.uint32))
    else:
        valid = ref["h64"] != T.NODATA
        assert np.array_equal(valid, h != T.NODATA)
        rel = np.abs(h[valid] - ref["h64"][valid]) / np.maximum(np.abs(ref["h64"][valid]), REL_FLOOR)
        assert rel.max() <= REL_TOL
    assert (ref["count"] > 0).mean() > 0.3


def test_space_time_device_buffers_and_zero_radius(ctx):
    """The device-buffer call equals the host one; radius 0 equals the 2-D
    tile of the points at the current time."""
    t = S.make_tile(31, 10000, 64, kind="clustered")
    times = np.random.default_rng(31).integers(2000, 2004, t.n).astype(np.int32)
    h_host = ctx.tile_run_time(t.params, t.X, t.Y, t.Z, times, 2001.0, 1.0)
    dX, dY, dZ, dT = ctx.upload(t.X), ctx.upload(t.Y), ctx.upload(t.Z), ctx.upload(times)
    dH = ctx.alloc((t.params.nrows, t.params.ncols), np.float32)
    ctx.tile_run_time_dev(t.params, dX, dY, dZ, dT, t.n, 2001.0, 1.0, dH)
    assert np.array_equal(dH.to_host().view(np.uint32), h_host.view(np.uint32))
    for a in (dX, dY, dZ, dT, dH):
        a.free()
    h0 = ctx.tile_run_time(t.params, t.X, t.Y, t.Z, times, 2001.0, 0.0)
    m = times == 2001
    h2, _, _ = ctx.tile_run(t.params, t.X[m], t.Y[m], t.Z[m])
    assert np.array_equal(h0.view(np.uint32), h2.view(np.uint32))


def test_space_time_empty_and_bucket_sweep(ctx):
    t = S.make_tile(41, 6000, 48, kind="uniform")
    times = np.random.default_rng(41).integers(2000, 2005, t.n).astype(np.int32)
    ref = O.run_time(t.params, t.X, t.Y, t.Z, times, 2002.0, 2.0)
    for bm, tc in ((1, 8), (3, 30), (5, 17)):
        p = dataclasses.replace(t.params, bucket_px=bm, tile_cells=tc)
        h = ctx.tile_run_time(p, t.X, t.Y, t.Z, times, 2002.0, 2.0)
        assert np.array_equal(h.view(np.uint32), ref["h_fixed"].view(np.uint32)), (bm, tc)
    e = ctx.tile_run_time(t.params, t.X[:0], t.Y[:0], t.Z[:0], times[:0], 2002.0, 2.0)
    assert (e == T.NODATA).all()


def test_space_time_sorted_ties_across_times(ctx):
    """Sorted points repeated two time units later at the same place: in the
    slice halfway between, both are at the same distance and the earlier
    point (the earlier time) wins, in the per-time layers as in the oracle;
    the random order of the same points (every slice searched) matches the
    oracle on its own order too."""
    t = S.make_tile(51, 6000, 40, kind="clustered")
    rng = np.random.default_rng(51)
    times = rng.integers(2000, 2004, t.n).astype(np.int32)
    k = rng.choice(t.n, 600, replace=False)
    X, Y = np.concatenate([t.X, t.X[k]]), np.concatenate([t.Y, t.Y[k]])
    Z = np.concatenate([t.Z, t.Z[k] + 777])
    Tm = np.concatenate([times, times[k] + 2])
    for order in ("sorted", "unsorted"):
        o = np.argsort(Tm, kind="stable") if order == "sorted" else rng.permutation(len(X))
        for cur, r in ((2002.0, 1.0), (2003.0, 2.0), (2002.5, 1.5)):
            h = ctx.tile_run_time(t.params, X[o], Y[o], Z[o], Tm[o], cur, r)
            ref = O.run_time(t.params, X[o], Y[o], Z[o], Tm[o], cur, r)
            assert np.array_equal(h.view(np.uint32), ref["h_fixed"].view(np.uint32)), (order, cur, r)
