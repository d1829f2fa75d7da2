"""The space-time slice cache on the GPU (performNNI's Delta algorithm,
nni.cpp:281-379 and 902-921): a caller steps cur_time through a
time-sorted stream, each call with the points of a window around cur_time
(as the nni.h layer reads them); slices already held are gathered without
a search.  Every cached call must equal the uncached call bit for bit, and
the uncached call the oracle; the hit and miss counts follow from the
slice times."""
import dataclasses

import numpy as np
import pytest

import pyoracle as O
import terranni_amd as T
import terranni_amd.synth as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _stream(seed, n, grid, alg, mode, t0=1996, t1=2010, dup=200, **kw):
    """A tile whose points are sorted by time (stable), with `dup` points
    repeated at the same place and time but another height, so that equal
    distances are decided by the earlier point."""
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
        # 5 slices in the first call, then one new slice per call
        assert c.stats() == (6 * 4, 5 + 6)


def test_cache_keys_tile_radius_and_slice_time(ctx):
    """Another radius, another tile or a clear empties the cache; slice times
    between integers are their own keys; going back in time reuses the
    slices still held."""
    p, X, Y, Z, Tm = _stream(21, 10000, 48, T.ALG_PLANES, T.ACC_FIXED64)
    with ctx.time_cache() as c:
        def run(q, cur, r):
            w = _window(Tm, cur, r)
            h = ctx.tile_run_time(q, X[w], Y[w], Z[w], Tm[w], cur, r, cache=c)
            assert _same(h, ctx.tile_run_time(q, X[w], Y[w], Z[w], Tm[w], cur, r)), (cur, r)

        run(p, 2001.0, 1.0)
        assert c.stats() == (0, 3)
        run(p, 2002.0, 1.0)             # 2001, 2002 held
        assert c.stats() == (2, 4)
        run(p, 2001.5, 1.0)             # 2000.5 .. 2002.5: none held
        assert c.stats() == (2, 7)
        run(p, 2001.5, 1.0)
        assert c.stats() == (5, 7)
        run(p, 2001.5, 1.5)             # another radius
        assert c.stats() == (5, 10)
        q = dataclasses.replace(p, origin_x=p.origin_x + 0.5)
        run(q, 2001.5, 1.5)             # another tile
        assert c.stats() == (5, 13)
        run(q, 2001.5, 1.5)
        assert c.stats() == (8, 13)
        c.clear()
        run(q, 2001.5, 1.5)
        assert c.stats() == (8, 16)
        # r = 0: one slice per call, one slot
        run(q, 2003.0, 0.0)
        run(q, 2003.0, 0.0)
        run(q, 2004.0, 0.0)
        run(q, 2003.0, 0.0)
        assert c.stats() == (9, 19)


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
        # 1999..2003, 2000..2004 (4 held), 2007..2011, 2008..2012 (4 held),
        # 2001..2005 (the slots hold 2008..2012)
        assert c.stats() == (8, 17)


def test_cache_from_another_context_is_refused(ctx):
    p, X, Y, Z, Tm = _stream(41, 2000, 24, T.ALG_PLANES, T.ACC_FIXED64, dup=0)
    with T.Context(0) as other:
        c = other.time_cache()
        with pytest.raises(T.TnniError) as e:
            ctx.tile_run_time(p, X, Y, Z, Tm, 2001.0, 1.0, cache=c)
        assert e.value.status == -1
        c.close()


def test_cache_with_slots_for_half_steps(ctx):
    """cur_time stepped by 0.5 alternates two phases of slice times; a cache
    of 2 x 5 slots keeps both, so every call after the first two searches
    one slice."""
    p, X, Y, Z, Tm = _stream(51, 8000, 40, T.ALG_PLANES, T.ACC_FP32_2SUM)
    with ctx.time_cache(slots=10) as c, ctx.time_cache() as d:
        for k in range(8):
            cur = 2001.0 + 0.5 * k
            w = _window(Tm, cur, 2.0)
            h = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 2.0, cache=c)
            g = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 2.0, cache=d)
            u = ctx.tile_run_time(p, X[w], Y[w], Z[w], Tm[w], cur, 2.0)
            assert _same(h, u) and _same(g, u), cur
        assert c.stats() == (6 * 4, 10 + 6)
        # 5 slots: each phase evicts the other's slices
        assert d.stats() == (0, 40)
    with pytest.raises(T.TnniError):
        ctx.time_cache(slots=5000)
