"""Pins the oracle's space-time extension (performNNI with timeRadius > 0,
nni.cpp:902-936) without the reference, which cannot run here:

* closed form: when every point has the time curTime, each slice j sees the
  2-D sites shifted by the same time term j^2 and its queries by the same
  j^2, so every slice steals exactly the 2-D pairs: heights equal the 2-D
  heights and counts are (2*(int)r + 1) times the 2-D counts;
* r = 0 is the 2-D tile of the points whose time is curTime;
* the slice's labels by brute force equal the raster pass;
* the bisector of two sites one time unit apart in a slice sits where
  (x - a)^2 = (x - b)^2 + 1 (metres), found analytically.
"""
import ctypes

import numpy as np
import pytest

import pyoracle as O
import terranni_amd as T
import terranni_amd.synth as S


def tile(seed=5, n=3000, grid=48, kind="uniform", **kw):
    return S.make_tile(seed, n, grid, kind=kind, **kw)


@pytest.mark.parametrize("r,alg", [(1.0, T.ALG_PLANES), (2.0, T.ALG_PLANES), (1.5, T.ALG_CONES)])
def test_time_uniform_points_give_the_2d_surface(r, alg):
    t = tile(seed=11, algorithm=alg)
    times = np.full(t.n, 2002, np.int32)
    ref2 = O.run(t.params, t.X, t.Y, t.Z, s0=True)
    ref3 = O.run_time(t.params, t.X, t.Y, t.Z, times, 2002.0, r, s0=True)
    nsl = 2 * int(r) + 1
    assert np.array_equal(ref3["h_fixed"].view(np.uint32), ref2["h_fixed"].view(np.uint32))
    assert np.array_equal(ref3["count"], nsl * ref2["count"])
    assert ref3["S"] == nsl * ref2["S"]
    assert (ref2["count"] > 0).mean() > 0.5


def test_closed_form_catches_a_dropped_query_time_term():
    """The same sums with the queries' j^2 left out (a plausible slip)
    differ from the 2-D surface, so the closed form above pins that term."""
    t = tile(seed=12)
    times = np.full(t.n, 7, np.int32)
    d, f, xu, yu, keep = O._prep(t.params, t.X, t.Y)
    n = xu.shape[0]
    acc = np.zeros(t.params.nrows * t.params.ncols, O.ACC)
    ks = np.empty(n, np.uint8)
    dt2 = np.empty(n, np.uint32)
    Z = np.ascontiguousarray(t.Z, np.int32)
    for j in (-1, 0, 1):
        O.lib().orc_slice_sites(ctypes.byref(f), O._p(times), O._p(keep), n, float(7 + j), 1.0,
                                O._p(ks), O._p(dt2))
        d2 = np.empty((f.H, f.W), np.uint32)
        lab = np.empty((f.H, f.W), np.int32)
        O.lib().orc_labels_raster_t(ctypes.byref(f), O._p(xu), O._p(yu), O._p(ks), O._p(dt2), n,
                                    O._p(d2), O._p(lab), 0)
        O.lib().orc_heights_accumulate(ctypes.byref(d), ctypes.byref(f), O._p(d2), O._p(lab),
                                       O._p(Z), 0, O._p(acc), None, 0)  # j^2 dropped
    hf = np.empty((t.params.nrows, t.params.ncols), np.float32)
    O.lib().orc_heights_finish(ctypes.byref(d), O._p(acc), O._p(hf), None, None, None)
    ref2 = O.run(t.params, t.X, t.Y, t.Z)
    assert not np.array_equal(hf, ref2["h_fixed"])


@pytest.mark.parametrize("seed", [1, 2])
def test_zero_radius_is_the_2d_tile_of_the_current_time(seed):
    t = tile(seed=seed, n=4000, kind="clustered")
    rng = np.random.default_rng(seed)
    times = rng.integers(2000, 2004, t.n).astype(np.int32)
    ref3 = O.run_time(t.params, t.X, t.Y, t.Z, times, 2001.0, 0.0)
    m = times == 2001
    ref2 = O.run(t.params, t.X[m], t.Y[m], t.Z[m])
    assert np.array_equal(ref3["h_fixed"].view(np.uint32), ref2["h_fixed"].view(np.uint32))
    assert np.array_equal(ref3["count"], ref2["count"])


@pytest.mark.parametrize("alg", [T.ALG_PLANES, T.ALG_CONES])
def test_slice_labels_brute_force_equals_raster(alg):
    t = tile(seed=21, n=300, grid=24, algorithm=alg)
    rng = np.random.default_rng(3)
    times = rng.integers(0, 5, t.n).astype(np.int32)
    a = O.run_time(t.params, t.X, t.Y, t.Z, times, 2.0, 2.0, keep_slices=True)
    b = O.run_time(t.params, t.X, t.Y, t.Z, times, 2.0, 2.0, keep_slices=True, brute=True)
    for (la, da), (lb, db) in zip(a["slices"], b["slices"]):
        assert np.array_equal(la, lb) and np.array_equal(da, db)
    assert np.array_equal(a["h_fixed"].view(np.uint32), b["h_fixed"].view(np.uint32))


def test_time_term_moves_the_bisector_analytically():
    """Sites a = (0.6, 0.6) m at time 10 and b = (2.6, 0.6) m at time 11; in
    the slice t = 10 a pixel centre x on the row y = 0.6 takes a iff
    (x - 0.6)^2 <= (x - 2.6)^2 + 1, i.e. x <= 1.85 (equality: the lower
    index, a).  0.2 m pixels, centres at -2.8 + 0.2 k."""
    p = T.TileParams(nrows=4, ncols=4, cell_size=1.0, scale=5, site_radius=5.0,
                     las_scale=(0.01, 0.01, 0.01))
    X = np.array([60, 260], np.int32)
    Y = np.array([60, 60], np.int32)
    Z = np.array([100, 300], np.int32)
    times = np.array([10, 11], np.int32)
    out = O.run_time(p, X, Y, Z, times, 10.0, 1.0, keep_slices=True)
    f = O.frame(p)
    lab, _ = out["slices"][1]  # j = 0: t = 10
    ps, upm = f["ps_u"], f["upm"]
    yrow = [l for l in range(f["H"]) if f["y0"] + l * ps == 60]
    assert yrow, "a pixel row through y = 0.6 m"
    l = yrow[0]
    checked = 0
    for k in range(f["W"]):
        x = (f["x0"] + k * ps) / upm
        if abs(x - 0.6) * upm > f["Rs"] or abs(x - 2.6) * upm > f["Rs"]:
            continue
        want = 0 if (x - 0.6) ** 2 <= (x - 2.6) ** 2 + 1.0 else 1
        assert lab[l, k] == want, (x, lab[l, k])
        checked += 1
    assert checked >= 8
    # the slice t = 9 holds a only (b is 2 > r away); t = 11 holds both
    lab9, _ = out["slices"][0]
    assert set(np.unique(lab9)) <= {-1, 0}
    lab11, _ = out["slices"][2]
    assert {0, 1} <= set(np.unique(lab11))
