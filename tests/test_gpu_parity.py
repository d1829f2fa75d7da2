"""GPU parity: labels bit-exact, heights bit-exact to the oracle's engine
formula and within 1e-5 relative of its FP64 mean, on seeded tiles."""
import dataclasses

import numpy as np
import pytest

import terranni_amd as T
import terranni_amd.synth as S
from helpers import check_tile

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("seed,n,grid,cs,scale,kind", [
    (1, 500, 32, 0.5, 1, "uniform"),
    (2, 3000, 64, 0.5, 1, "clustered"),
    (3, 2000, 40, 2.0, 5, "uniform"),
    (4, 800, 33, 1.0, 2, "uniform"),
    (5, 20000, 128, 0.5, 1, "clustered"),
    (6, 5000, 50, 2.0, 3, "clustered"),
])
def test_small_tiles(ctx, seed, n, grid, cs, scale, kind):
    _say("small tile", seed, n, grid, cs, scale, kind)
    t = S.make_tile(seed, n, grid, cell_size=cs, scale=scale, kind=kind)
    check_tile(ctx, t.params, t.X, t.Y, t.Z)


@pytest.mark.parametrize("mode", [T.ACC_FIXED64, T.ACC_FP32_2SUM, T.ACC_S0])
@pytest.mark.parametrize("alg", [T.ALG_PLANES, T.ALG_CONES])
def test_modes(ctx, mode, alg):
    _say("modes", mode, alg)
    t = S.make_tile(11, 6000, 96, cell_size=0.5, scale=1, kind="clustered")
    p = dataclasses.replace(t.params, accum_mode=mode, algorithm=alg, point_min=-300,
                            multiplier=10000)
    check_tile(ctx, p, t.X, t.Y, t.Z, s0=(mode == T.ACC_S0))


@pytest.mark.parametrize("form", [T.FORM_AUTO, T.FORM_FUSED, T.FORM_SPLIT, T.FORM_SCATTER])
def test_uniform_100k_512(ctx, form):
    """The reference's CPU case (scale 5: a 17-pixel query window) in every
    form; automatic picks the split form here."""
    _say("uniform 100k, 512^2 at 2 m, scale 5, form", form)
    t = S.workload_uniform_100k()
    p = dataclasses.replace(t.params, form=form)
    check_tile(ctx, p, t.X, t.Y, t.Z)


@pytest.mark.parametrize("mode", [T.ACC_FIXED64, T.ACC_FP32_2SUM, T.ACC_S0])
@pytest.mark.parametrize("alg", [T.ALG_PLANES, T.ALG_CONES])
def test_split_form_without_label_outputs_equals_fused(ctx, mode, alg):
    """Heights only (the split form's rasters then live in the engine's
    scratch), on host and device buffers: bit-equal to the fused form, in
    every height mode and region shape (the gather sums the same pixels in
    the same order), and so is the scatter form in the integer modes."""
    t = S.make_tile(12, 30000, 120, cell_size=2.0, scale=4, kind="clustered")
    base = dataclasses.replace(t.params, accum_mode=mode, algorithm=alg, point_min=-300, multiplier=10000)
    hs = []
    forms = (T.FORM_FUSED, T.FORM_SPLIT) + ((T.FORM_SCATTER,) if mode != T.ACC_FP32_2SUM else ())
    for form in forms:
        p = dataclasses.replace(base, form=form)
        h, _, _ = ctx.tile_run(p, t.X, t.Y, t.Z)
        dX, dY, dZ = ctx.upload(t.X), ctx.upload(t.Y), ctx.upload(t.Z)
        dh = ctx.alloc((p.nrows, p.ncols), np.float32)
        ctx.tile_run_dev(p, dX, dY, dZ, t.n, dh)
        hd = dh.to_host()
        for a in (dX, dY, dZ, dh):
            a.free()
        assert np.array_equal(h.view(np.uint32), hd.view(np.uint32))
        hs.append(h)
    for h in hs[1:]:
        assert np.array_equal(hs[0].view(np.uint32), h.view(np.uint32))
    check_tile(ctx, dataclasses.replace(base, form=T.FORM_SPLIT), t.X, t.Y, t.Z,
               s0=mode == T.ACC_S0)


@pytest.mark.parametrize("scale", [1, 2, 3, 5, 7])
@pytest.mark.parametrize("kind", ["clustered", "uniform"])
def test_scatter_form_matches_the_oracle(ctx, scale, kind):
    """The scatter form (each pixel spread over the cells its nearest-site
    distance reaches, lanes that hit one cell summed across the wavefront
    first when the scale puts several pixels on a cell) at scales 1-7,
    with labels: labels, squared distances and int64 heights bit-equal to
    the oracle, S0 heights too, cones as well as planes."""
    _say("scatter", scale, kind)
    t = S.make_tile(40 + scale, 20000, 90, cell_size=1.0, scale=scale, kind=kind)
    for mode, alg in ((T.ACC_FIXED64, T.ALG_PLANES), (T.ACC_S0, T.ALG_CONES)):
        p = dataclasses.replace(t.params, accum_mode=mode, algorithm=alg, point_min=-300, multiplier=10000,
                                form=T.FORM_SCATTER)
        check_tile(ctx, p, t.X, t.Y, t.Z, s0=mode == T.ACC_S0)


def test_scatter_form_refuses_fp32_two_sum(ctx):
    t = S.make_tile(3, 500, 32, cell_size=0.5, scale=1, kind="uniform")
    p = dataclasses.replace(t.params, accum_mode=T.ACC_FP32_2SUM, form=T.FORM_SCATTER)
    with pytest.raises(T.TnniError, match="scatter form needs an integer"):
        ctx.tile_run(p, t.X, t.Y, t.Z)


@pytest.mark.parametrize("bucket", [0, 1, 2, 3, 5, 8])
def test_bucket_size_does_not_change_results(ctx, bucket):
    _say("bucket", bucket)
    t = S.make_tile(21, 8000, 80, cell_size=0.5, scale=1, kind="clustered")
    p = dataclasses.replace(t.params, bucket_px=bucket)
    check_tile(ctx, p, t.X, t.Y, t.Z)


@pytest.mark.parametrize("tc,scale,po", [(1, 1, -1), (3, 1, -1), (16, 1, -1), (48, 1, -1),
                                         (1, 3, 0), (2, 3, 0), (5, 3, -1)])
def test_cell_tile_size_does_not_change_results(ctx, tc, scale, po):
    """Scale 3 without a border: the query halo of a 1- or 2-cell item
    (pixels -4..6 of its cell) reaches past its neighbour, so the items at
    the raster edge are clipped unevenly and the widest item (11 pixels at
    tc 1) is neither of the first two nor the last (10, 7, 7)."""
    _say("tile cells", tc, "scale", scale, "pixel offset", po)
    t = S.make_tile(22, 4000, 70, cell_size=2.0, scale=scale, kind="uniform")
    p = dataclasses.replace(t.params, tile_cells=tc, pixel_offset=po)
    check_tile(ctx, p, t.X, t.Y, t.Z)


@pytest.mark.parametrize("kind,bucket", [("clustered", 0), ("uniform", 0), ("uniform", 3)])
def test_dense_tile_searches_ring_by_ring(ctx, kind, bucket):
    """About 38 points a pixel: even 1-pixel buckets hold dozens of sites, so
    the searches go ring by ring (the pixel's own bucket usually settles
    them) instead of over whole squares of buckets; same labels either way."""
    _say("dense", kind, "bucket", bucket)
    t = S.make_tile(31, 200000, 64, cell_size=0.5, scale=1, kind=kind)
    p = dataclasses.replace(t.params, bucket_px=bucket)
    check_tile(ctx, p, t.X, t.Y, t.Z)


def test_empty_tile(ctx):
    _say("empty tile")
    t = S.make_tile(1, 10, 16)
    e = np.zeros(0, np.int32)
    h, lab, d2 = ctx.tile_run(t.params, e, e, e, labels=True)
    assert (h == T.NODATA).all()
    assert (lab == -1).all() and (d2 == T.NO_D2).all()


def test_points_outside_tile(ctx):
    _say("points far outside")
    t = S.make_tile(1, 1000, 16)
    X = t.X + 10**7
    h, lab, _ = ctx.tile_run(t.params, X, t.Y, t.Z, labels=True)
    assert (h == T.NODATA).all() and (lab == -1).all()


def test_single_site(ctx):
    _say("single site")
    t = S.make_tile(1, 1, 32, cell_size=0.5, scale=1)
    h, ref = check_tile(ctx, t.params, t.X, t.Y, t.Z)
    valid = h != T.NODATA
    assert valid.any()
    z = t.Z[0] * 0.01
    assert np.allclose(h[valid], z, rtol=1e-6)


def test_exact_duplicates_take_lowest_index(ctx):
    _say("duplicates")
    t = S.make_tile(3, 400, 24, cell_size=0.5, scale=1)
    X = np.concatenate([t.X, t.X[::-1]])
    Y = np.concatenate([t.Y, t.Y[::-1]])
    Z = np.concatenate([t.Z, t.Z[::-1] + 100])
    h, lab, _ = ctx.tile_run(t.params, X, Y, Z, labels=True)
    assert lab.max() < t.n  # the originals (lower indices) always win
    check_tile(ctx, t.params, X, Y, Z)


@pytest.mark.parametrize("seed,grid,cs,scale,lattice,bucket", [
    (11, 512, 0.5, 1, True, 0),
    (12, 512, 0.5, 1, False, 0),
    (13, 256, 1.0, 2, True, 3),
    (14, 384, 0.5, 1, False, 2),
])
def test_cluster_beside_sparse_stretch(ctx, seed, grid, cs, scale, lattice, bucket):
    """A few pixels between sparse sites whose search squares reach a dense
    cluster: the long searches that lane groups share.  A lattice cluster
    in shuffled order puts equal distances on different lanes of a group,
    so the lower index must still win."""
    _say("cluster beside sparse stretch", seed, grid, cs, scale, lattice, bucket)
    rng = np.random.default_rng(seed)
    ext = grid * cs
    border = 5 * cs + 5.0
    c = rng.uniform(0.3 * ext, 0.7 * ext, 2)
    if lattice:
        g = np.arange(-70, 70) * 0.03
        cx, cy = (a.ravel() for a in np.meshgrid(c[0] + g, c[1] + g))
    else:
        cx, cy = c[0] + rng.normal(0, 1.0, 20000), c[1] + rng.normal(0, 1.0, 20000)
    sx, sy = rng.uniform(-border, ext + border, (2, 60))
    x, y = np.concatenate([cx, sx]), np.concatenate([cy, sy])
    o = rng.permutation(x.shape[0])
    x, y = x[o], y[o]
    z = rng.normal(100.0, 5.0, x.shape[0])
    X, Y, Z = S._to_las(x, y, z, (0.0, 0.0))
    p = S._params(grid, grid, cs, scale, (0.0, 0.0), bucket_px=bucket)
    check_tile(ctx, p, X, Y, Z)


def test_single_4m_tile_labels_and_heights(ctx):
    """The 4096 x 4096 tile with 4M clustered points, labels emitted."""
    _say("4M-point 4096^2 tile")
    t = S.workload_single_4m()
    h, ref = check_tile(ctx, t.params, t.X, t.Y, t.Z)
    _say("valid cells", int((h != T.NODATA).sum()), "stolen pairs", ref["S"])
