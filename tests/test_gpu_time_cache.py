"""The space-time cache on the GPU (what performNNI's Delta algorithm saves,
nni.cpp:281-379 and 902-921): a caller steps cur_time through a
time-sorted stream, each call with the points of a window around cur_time
(as the nni.h layer reads them); the cache keeps per point time the 2-D
Voronoi layer of that time's points, each slice is combined from the
layers within its radius, and only missing layers are searched.  Every
cached call must equal the uncached call (which searches every slice) bit
for bit, and the uncached call the oracle; the layer hit and miss counts
follow from the times the calls reach."""
import dataclasses

import numpy as np
import pytest

import pyoracle as O
import terranni_amd as T
import terranni_amd.synth as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _stream(seed, n, grid, alg, mode, t0=1996, t1=2010, dup=200, xdup=0, **kw):
    """A tile whose points are sorted by time (stable), with `dup` points
    repeated at the same place and time but another height, and `xdup`
    repeated at the same place two time units later, so that equal
    distances are decided by the earlier point (within a time and, for a
    slice halfway between, across times)."""
    t = S.make_tile(seed, n, grid, kind="clustered", algorithm=alg, accum_mode=mode,
                    point_min=-60, multiplier=1000, **kw)
    rng = np.random.default_rng(seed)
    times = rng.integers(t0, t1, t.n).astype(np.int32)
    X, Y, Z = t.X.copy(), t.Y.copy(), t.Z.copy()
    if dup:
        k = rng.choice(t.n, dup, replace=False)
        X, Y = np.concatenate([X, X[k]]), np.concatenate([Y, Y[k]])
        Z = np.concatenate([Z, Z[k] + rng.integers(1, 500, dup).astype(np.int32)])
        times = np.concatenate([times, times[k]])
    if xdup:
        k = rng.choice(t.n, xdup, replace=False)
        X, Y = np.concatenate([X, X[k]]), np.concatenate([Y, Y[k]])
        Z = np.concatenate([Z, Z[k] - rng.integers(1, 500, xdup).astype(np.int32)])
        times = np.concatenate([times, times[k] + 2])
    o = np.argsort(times, kind="stable")
    return t.params, X[o], Y[o], Z[o], times[o]


def _window(T_, cur, r):
    """The points the nni.h layer reads for cur_time: |T - cur| <= 2r + 1
    (every site of every slice of the call)."""
    w = 2.0 * r + 1.0
    m = np.abs(T_.astype(np.float64) - cur) <= w
    return np.flatnonzero(m)


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("alg", [T.ALG_PLANES, T.ALG_CONES])
@pytest.mark.parametrize("mode", [T.ACC_FIXED64, T.ACC_FP32_2SUM, T.ACC_S0])
def test_cached_slices_equal_uncached(ctx, alg, mode):
    r = 2.0
    p, X, Y, Z, Tm = _stream(11 + alg * 3 + mode, 12000, 56, alg, mode)
    _say("cache", alg, mode, p.nrows, len(X))
    with ctx.time_cache() as c:
        for k, cur in enumerate(range(2000, 2007)):
            w = _window(Tm, cur, r)
            h = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], float(cur), r, cache=c)
            u = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], float(cur), r)
            assert _same(h, u), cur
            if k == 6:
                ref = O.run_time(p, X[w], Y[w], Z[w], Tm[w], float(cur), r, s0=mode == T.ACC_S0)
                if mode == T.ACC_FIXED64:
                    assert _same(h, ref["h_fixed"])
                elif mode == T.ACC_S0:
                    assert _same(h, ref["h_s0"])
        # layers of the times within 4 of cur_time: 9 in the first call,
        # then one new layer per call
        assert c.stats() == (6 * 8, 9 + 6)


def test_cache_keys_tile_and_point_time(ctx):
    """Layers are keyed by point time: slices between integers and other
    radii reuse them; another tile or a clear empties the cache; a slice
    with no point time within its radius has no sites."""
    p, X, Y, Z, Tm = _stream(21, 10000, 48, T.ALG_PLANES, T.ACC_FIXED64)
    with ctx.time_cache() as c:
        def run(q, cur, r):
            w = _window(Tm, cur, r)
            h = ctx.tile_run_time(q, X[w], Y[w], Z[w], Tm[w], cur, r, cache=c)
            assert _same(h, ctx.tile_run_time(q, X[w], Y[w], Z[w], Tm[w], cur, r)), (cur, r)

        run(p, 2001.0, 1.0)             # times 1999 .. 2003
        assert c.stats() == (0, 5)
        run(p, 2002.0, 1.0)             # 2000 .. 2004: 2004 is new
        assert c.stats() == (4, 6)
        run(p, 2001.5, 1.0)             # slices 2000.5 .. 2002.5: times 2000 .. 2003
        assert c.stats() == (8, 6)
        run(p, 2001.5, 1.5)             # another radius: 1999 .. 2004
        assert c.stats() == (13, 7)
        q = dataclasses.replace(p, origin_x=p.origin_x + 0.5)
        run(q, 2001.5, 1.5)             # another tile
        assert c.stats() == (13, 13)
        run(q, 2001.5, 1.5)
        assert c.stats() == (19, 13)
        c.clear()
        run(q, 2001.5, 1.5)
        assert c.stats() == (19, 19)
        run(q, 2003.0, 0.0)             # r = 0: the one time, held
        assert c.stats() == (20, 19)
        run(q, 2003.5, 0.0)             # no point time at the slice: no sites
        assert c.stats() == (20, 19)


def test_cache_device_buffers_and_empty_windows(ctx):
    """The device-buffer call through a cache, windows with no points at
    all, and a tile at scale 3 with a wide query radius."""
    p, X, Y, Z, Tm = _stream(31, 8000, 40, T.ALG_CONES, T.ACC_FIXED64, t0=2000, t1=2004,
                             scale=3, query_radius=3.0)
    with ctx.time_cache() as c:
        for cur in (2001.0, 2002.0, 2009.0, 2010.0, 2003.0):
            w = _window(Tm, cur, 2.0)
            n = len(w)
            dX, dY, dZ, dT = (ctx.upload(np.ascontiguousarray(a[w])) if n else ctx.alloc((1,), np.int32)
                              for a in (X, Y, Z, Tm))
            dH = ctx.alloc((p.nrows, p.ncols), np.float32)
            ctx.tile_run_time_dev(p, dX, dY, dZ, dT, n, cur, 2.0, dH, cache=c)
            u = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 2.0)
            assert _same(dH.to_host(), u), cur
            if n == 0:
                assert (u == T.NODATA).all()
            for a in (dX, dY, dZ, dT, dH):
                a.free()
        # times 1997..2005, 1998..2006 (8 held), 2005..2013 (2 held),
        # 2006..2014 (8 held), 1999..2007 (2 held; 9 slots)
        assert c.stats() == (20, 25)


def test_cache_from_another_context_is_refused(ctx):
    p, X, Y, Z, Tm = _stream(41, 2000, 24, T.ALG_PLANES, T.ACC_FIXED64, dup=0)
    with T.Context(0) as other:
        c = other.time_cache()
        with pytest.raises(T.TnniError) as e:
            ctx.tile_run_time(p, X, Y, Z, Tm, 2001.0, 1.0, cache=c)
        assert e.value.status == -1
        c.close()


def test_cache_half_steps_reuse_layers(ctx):
    """cur_time stepped by 0.5: the slices of integer and half-integer
    cur_time reach the same point times, so every call after the first
    searches at most one layer, whatever the slot count."""
    p, X, Y, Z, Tm = _stream(51, 8000, 40, T.ALG_PLANES, T.ACC_FP32_2SUM)
    with ctx.time_cache(slots=10) as c, ctx.time_cache() as d:
        for k in range(8):
            cur = 2001.0 + 0.5 * k
            w = _window(Tm, cur, 2.0)
            h = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 2.0, cache=c)
            g = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 2.0, cache=d)
            u = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 2.0)
            assert _same(h, u) and _same(g, u), cur
        # times 1997..2005, 1998..2005, 1998..2006, 1999..2006, ...
        assert c.stats() == (56, 12)
        assert d.stats() == (56, 12)
    with pytest.raises(T.TnniError):
        ctx.time_cache(slots=5000)


def test_cache_ties_across_times_and_unsorted_points(ctx):
    """Points repeated two time units later at the same place tie with the
    original in the slice between them (equal plane and time distances):
    the earlier point wins, as in the slice search.  Points not sorted by
    time run without the cache (its counts stay) and still give the
    uncached result."""
    for alg in (T.ALG_PLANES, T.ALG_CONES):
        p, X, Y, Z, Tm = _stream(61 + alg, 6000, 40, alg, T.ACC_FIXED64, t0=2000, t1=2006,
                                 dup=50, xdup=400)
        with ctx.time_cache() as c:
            for cur in (2002.0, 2003.0, 2004.0):
                w = _window(Tm, cur, 1.0)
                h = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 1.0, cache=c)
                assert _same(h, ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 1.0)), cur
                ref = O.run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 1.0)
                assert _same(h, ref["h_fixed"]), cur
            before = c.stats()
            rng = np.random.default_rng(5)
            o = rng.permutation(len(X))
            h = ctx.tile_run_time(p, X[o], Y[o], Z[o], Tm[o], 2003.0, 1.0, cache=c)
            assert _same(h, ctx.tile_run_time(p, X[o], Y[o], Z[o], Tm[o], 2003.0, 1.0))
            assert c.stats() == before


def test_many_point_times_fall_back_to_slice_search(ctx):
    """A call reaching more than 64 point times (r = 20 over 60 years: the 81
    times within 40 of cur_time) searches every slice, with or without a
    cache (whose counts stay); r = 10 (41 times) uses the layers.  Both
    equal the oracle."""
    p, X, Y, Z, Tm = _stream(71, 3000, 24, T.ALG_PLANES, T.ACC_FIXED64, t0=1970, t1=2030, dup=0)
    with ctx.time_cache() as c:
        for r, layered in ((20.0, False), (10.0, True)):
            before = c.stats()
            h = ctx.tile_run_time(p, X, Y, Z, Tm, 2000.0, r, cache=c)
            u = ctx.tile_run_time(p, X, Y, Z, Tm, 2000.0, r)
            ref = O.run_time(p, X, Y, Z, Tm, 2000.0, r)
            assert _same(h, u) and _same(h, ref["h_fixed"]), r
            hits, misses = c.stats()
            assert (misses > before[1]) == layered, (r, before, (hits, misses))
