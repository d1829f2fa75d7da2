"""Seeded sweep of tile geometries on the GPU against the oracle.

Each case draws, from its own seed: cell size, scale, site and query radius,
region shape (planes / cones), height mode, bucket and cell-tile sizes,
pixel offset, grid shape (rows may differ from columns), point count and
distribution (uniform, clustered, a few, none), and for a quarter of the
cases a space-time tile, run with its points in random time order and
sorted by time.  2-D cases run in every form (fused: labels with the query
halo and accumulation in one kernel; split: each pixel labelled once into a
raster, then a gather pass; scatter, in the integer height modes: the split
form's raster, then each pixel spread over the cells it reaches).  Labels and their squared distances must be
bit-equal to the oracle's, heights bit-equal to the exact formula (int64
fixed point, S0) or within 1e-5 relative of FP64 (FP32 two-sum), the
NODATA sets equal.  A frame the fixed-point grid cannot hold must be
refused by both."""
import dataclasses

import numpy as np
import pytest

import pyoracle as O
import terranni_amd as T
import terranni_amd.synth as S
from helpers import REL_TOL, rel_err

pytestmark = pytest.mark.gpu

N_CASES = 160


def _say(*a):
    print(*a, flush=True)


def draw_case(k):
    rng = np.random.default_rng(1000 + k)
    cs = float(rng.choice([0.25, 0.5, 1.0, 1.5, 2.0, 3.0]))
    scale = int(rng.integers(1, 6))
    grid = int(rng.integers(6, 80))
    rows = grid if rng.random() < 0.6 else int(rng.integers(4, 90))
    n = int(rng.choice([0, 40, 3000, 20000, 60000]))
    kind = "clustered" if rng.random() < 0.5 else "uniform"
    p = dict(site_radius=float(rng.choice([1.0, 2.5, 5.0, 8.0])),
             algorithm=int(rng.integers(0, 2)),
             accum_mode=int(rng.choice([T.ACC_FIXED64, T.ACC_FP32_2SUM, T.ACC_S0])),
             bucket_px=int(rng.choice([0, 0, 1, 2, 3, 5, 8])),
             tile_cells=int(rng.choice([0, 0, 1, 4, 13, 30, 64])),
             pixel_offset=int(rng.choice([-1, -1, 0, 2])),
             point_min=-60, multiplier=1000)
    qr = rng.choice(["default", "small", "wide"])
    if qr == "small":
        p["query_radius"] = 0.6 * cs
    elif qr == "wide":
        p["query_radius"] = 1.7 * cs
    time_r = float(rng.choice([0.0, 1.0, 1.5, 2.0])) if rng.random() < 0.25 else None
    return dict(seed=k, cs=cs, scale=scale, grid=grid, rows=rows, n=n, kind=kind, params=p, time_r=time_r)


def check_heights(mode, h, ref):
    if mode == T.ACC_S0:
        assert np.array_equal(h.view(np.uint32), ref["h_s0"].view(np.uint32))
    elif mode == T.ACC_FIXED64:
        assert np.array_equal(h.view(np.uint32), ref["h_fixed"].view(np.uint32))
    else:
        emax, _, same = rel_err(h, ref["h64"])
        assert same
        assert emax <= REL_TOL, emax


def run_or_auto(run, p):
    """run(p); an explicit cell-tile size whose items outgrow a workgroup's
    LDS is refused (TNNI_E_GEOMETRY), and the automatic size runs instead."""
    try:
        return run(p), p
    except T.TnniError as e:
        if not (p.tile_cells > 0 and e.status == -2 and "tile_cells" in str(e)):
            raise
    p = dataclasses.replace(p, tile_cells=0)
    return run(p), p


@pytest.mark.parametrize("k", range(N_CASES))
def test_random_geometry_matches_oracle(ctx, k):
    c = draw_case(k)
    _say("sweep", c)
    t = S.make_tile(c["seed"], max(c["n"], 1), c["grid"], cell_size=c["cs"], scale=c["scale"],
                    kind=c["kind"], **c["params"])
    X, Y, Z = t.X[:c["n"]], t.Y[:c["n"]], t.Z[:c["n"]]
    p = dataclasses.replace(t.params, nrows=c["rows"])
    try:
        O.frame(p)
    except ValueError:
        with pytest.raises(T.TnniError):
            T.frame(p)
        return
    mode = p.accum_mode
    if c["time_r"] is None:
        ref, hs = None, []
        forms = (T.FORM_FUSED, T.FORM_SPLIT) + ((T.FORM_SCATTER,) if mode != T.ACC_FP32_2SUM else ())
        for form in forms:
            q = dataclasses.replace(p, form=form)
            (h, lab, d2), q = run_or_auto(lambda r: ctx.tile_run(r, X, Y, Z, labels=True), q)
            if ref is None:
                ref = O.run(q, X, Y, Z, s0=mode == T.ACC_S0)
            assert np.array_equal(lab, ref["labels"]), (form, int((lab != ref["labels"]).sum()))
            assert np.array_equal(d2, ref["d2"]), form
            check_heights(mode, h, ref)
            hs.append(h)
        # the forms sum the same pixels (in the same order: FP32 two-sum)
        for h in hs[1:]:
            assert np.array_equal(hs[0].view(np.uint32), h.view(np.uint32))
    else:
        # the points in random time order (every slice searched), then
        # sorted by time (per-time layers, slices combined from them)
        times = np.random.default_rng(c["seed"]).integers(1998, 2005, c["n"]).astype(np.int32)
        o = np.argsort(times, kind="stable")
        for Xs, Ys, Zs, Ts in ((X, Y, Z, times), (X[o], Y[o], Z[o], times[o])):
            h, p = run_or_auto(lambda q: ctx.tile_run_time(q, Xs, Ys, Zs, Ts, 2001.0, c["time_r"]), p)
            ref = O.run_time(p, Xs, Ys, Zs, Ts, 2001.0, c["time_r"], s0=mode == T.ACC_S0)
            check_heights(mode, h, ref)


def test_sweep_covers_both_search_modes_and_time():
    """The draw reaches every height mode and region shape, empty and large
    point sets, dense tiles (more than two sites per bucket: searches ring
    by ring) and sparse ones, and space-time slices."""
    cases = [draw_case(k) for k in range(N_CASES)]
    dense = 0
    for c in cases:
        f = T.frame(dataclasses.replace(S.make_tile(c["seed"], 1, c["grid"], cell_size=c["cs"],
                                                    scale=c["scale"], **c["params"]).params,
                                        nrows=c["rows"]))
        bm = c["params"]["bucket_px"] or max(1, round((f["W"] * f["H"] / max(c["n"], 1)) ** 0.5))
        dense += c["n"] * bm * bm > 2 * f["W"] * f["H"]
    assert 5 <= dense <= N_CASES - 5
    assert sum(c["time_r"] is not None for c in cases) >= 6
    assert any(c["n"] == 0 for c in cases) and any(c["n"] >= 20000 for c in cases)
    assert len({c["params"]["accum_mode"] for c in cases}) == 3
    assert len({c["params"]["algorithm"] for c in cases}) == 2


@pytest.mark.parametrize("alg", [T.ALG_PLANES, T.ALG_CONES])
def test_random_geometries_in_one_batch(ctx, alg):
    """The sweep's 2-D tiles of one region shape and height mode, all in one
    batched launch (so tiles of every scale, radius, bucket size and grid
    shape share the sort passes and the work queue): every tile's labels,
    d2 and heights equal the oracle's as when it runs alone.  Cell-tile
    sizes are left to the engine (one explicit size too large for the LDS
    would refuse the whole batch).  The tiles take the fused, split and (in
    the integer modes) scatter forms in turn, so tiles of every form share
    the launch."""
    cases = [draw_case(k) for k in range(N_CASES)]
    for mode in (T.ACC_FIXED64, T.ACC_FP32_2SUM, T.ACC_S0):
        group = []
        for c in cases:
            if c["time_r"] is not None or c["params"]["algorithm"] != alg or c["params"]["accum_mode"] != mode:
                continue
            t = S.make_tile(c["seed"], max(c["n"], 1), c["grid"], cell_size=c["cs"], scale=c["scale"],
                            kind=c["kind"], **c["params"])
            forms = (T.FORM_FUSED, T.FORM_SPLIT) + ((T.FORM_SCATTER,) if mode != T.ACC_FP32_2SUM else ())
            p = dataclasses.replace(t.params, nrows=c["rows"], tile_cells=0, form=forms[len(group) % len(forms)])
            group.append((p, t.X[:c["n"]], t.Y[:c["n"]], t.Z[:c["n"]]))
        _say("batch", alg, mode, len(group), "tiles")
        assert len(group) >= 3
        dev = []
        for p, X, Y, Z in group:
            f = T.frame(p)
            dev.append((ctx.upload(np.ascontiguousarray(X, np.int32)), ctx.upload(np.ascontiguousarray(Y, np.int32)),
                        ctx.upload(np.ascontiguousarray(Z, np.int32)), ctx.alloc((p.nrows, p.ncols), np.float32),
                        ctx.alloc((f["H"], f["W"]), np.int32), ctx.alloc((f["H"], f["W"]), np.uint32)))
        ctx.batch_run_dev([g[0] for g in group], [d[0] for d in dev], [d[1] for d in dev],
                          [d[2] for d in dev], [g[1].shape[0] for g in group], [d[3] for d in dev],
                          [d[4] for d in dev], [d[5] for d in dev])
        for (p, X, Y, Z), d in zip(group, dev):
            ref = O.run(p, X, Y, Z, s0=mode == T.ACC_S0)
            assert np.array_equal(d[4].to_host(), ref["labels"])
            assert np.array_equal(d[5].to_host(), ref["d2"])
            check_heights(mode, d[3].to_host(), ref)
            for a in d:
                a.free()
