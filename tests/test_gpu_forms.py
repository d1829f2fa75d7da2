"""Batched and served forms on the GPU: each tile's result is the same
alone, in a batch, replayed from a captured graph, and in a served mix."""
import numpy as np
import pytest

import terranni_amd as T
import terranni_amd.synth as S
from helpers import check_tile

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _mixed_tiles(k=10, seed=100):
    out = []
    for i in range(k):
        grid = [48, 64, 96, 130][i % 4]
        n = [500, 4000, 20000, 1][i % 4] * (1 + i % 3)
        kind = "clustered" if i % 2 else "uniform"
        out.append(S.make_tile(seed + i, n, grid, kind=kind, origin=(i * 300.0, 0.0)))
    return out


def _upload(ctx, tiles):
    dev = []
    for t in tiles:
        dev.append((ctx.upload(t.X), ctx.upload(t.Y), ctx.upload(t.Z),
                    ctx.alloc((t.params.nrows, t.params.ncols), np.float32)))
    return dev


def _free(dev):
    for row in dev:
        for a in row:
            a.free()


def test_batch_equals_one_by_one(ctx):
    _say("batch vs single")
    tiles = _mixed_tiles()
    singles = [ctx.tile_run(t.params, t.X, t.Y, t.Z)[0] for t in tiles]
    dev = _upload(ctx, tiles)
    ctx.batch_run_dev([t.params for t in tiles], [d[0] for d in dev], [d[1] for d in dev],
                      [d[2] for d in dev], [t.n for t in tiles], [d[3] for d in dev])
    for t, d, h1 in zip(tiles, dev, singles):
        assert np.array_equal(d[3].to_host().view(np.uint32), h1.view(np.uint32))
    _free(dev)


def test_batch_graph_replays_with_eager_work_between(ctx):
    _say("graph replays")
    tiles = _mixed_tiles(8, seed=200)
    ref = [ctx.tile_run(t.params, t.X, t.Y, t.Z)[0] for t in tiles]
    dev = _upload(ctx, tiles)
    other = S.make_tile(999, 3000, 40)
    for rep in range(4):
        for d in dev:
            d[3].free()
        dev = [(d[0], d[1], d[2], ctx.alloc(d[3].shape, np.float32)) for d in dev]
        ctx.batch_run_dev([t.params for t in tiles], [d[0] for d in dev], [d[1] for d in dev],
                          [d[2] for d in dev], [t.n for t in tiles], [d[3] for d in dev])
        for d, h1 in zip(dev, ref):
            assert np.array_equal(d[3].to_host().view(np.uint32), h1.view(np.uint32)), rep
        check_tile(ctx, other.params, other.X, other.Y, other.Z, labels=False)
    _free(dev)


def test_batch_with_labels_matches_oracle(ctx):
    _say("batch labels vs oracle")
    import pyoracle as O
    tiles = _mixed_tiles(6, seed=300)
    dev = _upload(ctx, tiles)
    labs = []
    for t in tiles:
        f = T.frame(t.params)
        labs.append(ctx.alloc((f["H"], f["W"]), np.int32))
    ctx.batch_run_dev([t.params for t in tiles], [d[0] for d in dev], [d[1] for d in dev],
                      [d[2] for d in dev], [t.n for t in tiles], [d[3] for d in dev], labs)
    for t, d, lab in zip(tiles, dev, labs):
        ref = O.run(t.params, t.X, t.Y, t.Z)
        assert np.array_equal(lab.to_host(), ref["labels"])
        assert np.array_equal(d[3].to_host().view(np.uint32), ref["h_fixed"].view(np.uint32))
        lab.free()
    _free(dev)


def test_served_alone_and_in_the_mix():
    _say("served alone vs mix")
    tiles = _mixed_tiles(16, seed=400)
    with T.Context(0, n_slots=3, max_batch_tiles=5) as c:
        alone = []
        for t in tiles:
            pend = c.submit(t.params, t.X, t.Y, t.Z, labels=True)
            alone.append(pend.wait())
        pends = [c.submit(t.params, t.X, t.Y, t.Z, labels=True) for t in tiles]
        mixed = [p.wait() for p in pends]
        st = c.serve_stats()
    _say("serve stats", st)
    assert st["completed"] == 32 and st["batches"] >= 4
    for (h1, l1, d1), (h2, l2, d2) in zip(alone, mixed):
        assert np.array_equal(h1.view(np.uint32), h2.view(np.uint32))
        assert np.array_equal(l1, l2) and np.array_equal(d1, d2)
    plain = T.Context(0)
    for t, (h1, l1, _) in zip(tiles, alone):
        h0, l0, _ = plain.tile_run(t.params, t.X, t.Y, t.Z, labels=True)
        assert np.array_equal(h0.view(np.uint32), h1.view(np.uint32))
        assert np.array_equal(l0, l1)
    plain.close()


def test_served_rejects_tile_over_budget():
    _say("served budget")
    t = S.make_tile(1, 20000, 256)
    with T.Context(0, hbm_budget=1 << 20) as c:
        with pytest.raises(T.TnniError) as e:
            c.submit(t.params, t.X, t.Y, t.Z)
        assert e.value.status == -6
        small = S.make_tile(2, 50, 8)
        h, _, _ = c.submit(small.params, small.X, small.Y, small.Z).wait()
        assert h.shape == (8, 8)


def test_served_admission_waits_for_memory(ctx):
    """A budget that admits about two tiles at once: every request still
    completes, with the same result as the plain form."""
    _say("served admission")
    tiles = [S.make_tile(500 + i, 30000, 128, kind="clustered") for i in range(12)]
    fp = T.footprint(tiles[0].params, 30000)
    with T.Context(0, n_slots=4, hbm_budget=int(4 * fp)) as c:
        pends = [c.submit(t.params, t.X, t.Y, t.Z) for t in tiles]
        outs = [p.wait()[0] for p in pends]
        st = c.serve_stats()
    _say("serve stats", st)
    assert st["completed"] == 12 and st["max_inflight_tiles"] <= 3
    for t, h in zip(tiles, outs):
        h0, _, _ = ctx.tile_run(t.params, t.X, t.Y, t.Z)
        assert np.array_equal(h0.view(np.uint32), h.view(np.uint32))


def test_served_device_resident_equals_plain(ctx):
    _say("served device-resident")
    tiles = _mixed_tiles(10, seed=600)
    with T.Context(0, n_slots=2, max_batch_tiles=4) as c:
        dev = [(c.upload(t.X), c.upload(t.Y), c.upload(t.Z),
                c.alloc((t.params.nrows, t.params.ncols), np.float32)) for t in tiles]
        pends = [c.submit_dev(t.params, d[0], d[1], d[2], t.n, d[3]) for t, d in zip(tiles, dev)]
        for p in pends:
            p.wait()
        for t, d in zip(tiles, dev):
            h0, _, _ = ctx.tile_run(t.params, t.X, t.Y, t.Z)
            assert np.array_equal(d[3].to_host().view(np.uint32), h0.view(np.uint32))


def test_autotune_keeps_results_and_table_round_trips():
    """Tuned bucket and cell-tile sizes change the work split, never a
    result: every tile is bit-identical before and after tuning, in the
    batched and the plain form; the table reads back, clears and reloads."""
    tiles = _mixed_tiles(8, seed=300)
    with T.Context(0) as c:
        dev = _upload(c, tiles)
        P = [t.params for t in tiles]
        cols = ([d[0] for d in dev], [d[1] for d in dev], [d[2] for d in dev])
        ns = [t.n for t in tiles]
        hs = [d[3] for d in dev]
        c.batch_run_dev(P, *cols, ns, hs)
        ref = [h.to_host() for h in hs]
        n_classes = 0
        for batched in (True, False):
            _say(f"autotune batched={batched}")
            ent = c.autotune(P, *cols, ns, batched=batched, reps=1)
            want = {T.shape_class(t.params, t.n, batched) for t in tiles}
            assert {e["shape_class"] for e in ent} == want
            for e in ent:
                assert e["candidates"] >= 1
                assert e["ms_tuned"] <= e["ms_default"]
                assert e["bucket_px"] >= 1 and 1 <= e["tile_cells"] <= 64 and e["wg_per_cu"] >= 0
            n_classes += len(want)
        c.batch_run_dev(P, *cols, ns, hs)
        for r, h in zip(ref, hs):
            assert np.array_equal(r.view(np.uint32), h.to_host().view(np.uint32))
        for t, d, r in zip(tiles, dev, ref):
            c.tile_run_dev(t.params, d[0], d[1], d[2], t.n, d[3])
            assert np.array_equal(r.view(np.uint32), d[3].to_host().view(np.uint32))
        tab = c.tune_table()
        assert len(tab) == n_classes
        c.tune_clear()
        assert c.tune_table() == []
        c.tune_load(tab)
        key = lambda e: e["shape_class"]  # noqa: E731
        assert sorted(c.tune_table(), key=key) == sorted(tab, key=key)
        _free(dev)
