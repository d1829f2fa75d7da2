"""The allocation-free entry point (tnni_tile_run_dev_noalloc, through the C
ABI) against tnni_tile_run_dev on the same device buffers: heights, labels
and squared distances bit for bit, on two calls in a row, the second with a
smaller tile (the fixed scratch keeps the first call's contents), then a
tile too large for the scratch (refused before any launch) and a further
call that is still right."""
import numpy as np
import pytest

import terranni_amd as T
import terranni_amd.synth as S

pytestmark = pytest.mark.gpu


def _outputs(ctx, t):
    f = T.frame(t.params)
    return (ctx.alloc((t.params.nrows, t.params.ncols), np.float32),
            ctx.alloc((f["H"], f["W"]), np.int32), ctx.alloc((f["H"], f["W"]), np.uint32))


def _run_both(ctx, t):
    dX, dY, dZ = ctx.upload(t.X), ctx.upload(t.Y), ctx.upload(t.Z)
    a, b = _outputs(ctx, t), _outputs(ctx, t)
    print(f"tile {t.params.nrows}x{t.params.ncols}, {t.n} points", flush=True)
    ctx.tile_run_dev_noalloc(t.params, dX, dY, dZ, t.n, *a)
    ctx.tile_run_dev(t.params, dX, dY, dZ, t.n, *b)
    ha, la, da = (x.to_host() for x in a)
    hb, lb, db = (x.to_host() for x in b)
    for x in (dX, dY, dZ, *a, *b):
        x.free()
    return ha, la, da, hb, lb, db


def test_noalloc_equals_tile_run_dev_on_two_calls():
    with T.Context(0) as ctx:
        big = S.make_tile(31, 600_000, 1024, cell_size=0.5, kind="clustered")
        small = S.make_tile(32, 30_000, 200, cell_size=0.5, kind="uniform")
        for t in (big, small):
            ha, la, da, hb, lb, db = _run_both(ctx, t)
            assert np.array_equal(la, lb), f"{int((la != lb).sum())} labels differ"
            assert np.array_equal(da, db)
            assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
            assert (ha != T.NODATA).mean() > 0.5


def test_noalloc_split_form_equals_tile_run_dev():
    """A scale-5 tile takes the split form (its value and squared-distance
    rasters then also come from the fixed scratch): equal to
    tnni_tile_run_dev, after a larger fused call."""
    with T.Context(0) as ctx:
        fused = S.make_tile(34, 200_000, 512, cell_size=0.5, kind="clustered")
        split = S.workload_uniform_100k()
        for t in (fused, split):
            ha, la, da, hb, lb, db = _run_both(ctx, t)
            assert np.array_equal(la, lb) and np.array_equal(da, db)
            assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))


def test_noalloc_refuses_a_tile_over_its_scratch():
    with T.Context(0) as ctx:
        n = 5_000_000
        p = T.TileParams(nrows=256, ncols=256, cell_size=0.5, scale=1)
        z = np.zeros(n, np.int32)
        dX, dY, dZ = ctx.upload(z), ctx.upload(z), ctx.upload(z)
        dh = ctx.alloc((256, 256), np.float32)
        with pytest.raises(T.TnniError) as e:
            ctx.tile_run_dev_noalloc(p, dX, dY, dZ, n, dh)
        assert e.value.status == -6 and "bytes of scratch" in str(e.value)
        for x in (dX, dY, dZ, dh):
            x.free()
        # the next call is unaffected
        t = S.make_tile(33, 20_000, 128, cell_size=0.5, kind="clustered")
        ha, la, da, hb, lb, db = _run_both(ctx, t)
        assert np.array_equal(la, lb) and np.array_equal(da, db)
        assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
