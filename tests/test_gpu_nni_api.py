"""The reference's C++ API (nni.h) end to end on the GPU: a C++ program
feeds a time-sorted TPIE stream of Point4D to nni::performNNI slice by slice,
as terranni.cpp's binRun does, and the OutPoints it appends are compared
with the oracle on the same (quantised) points."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
import terranni_amd as T
import terranni_amd.synth as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "nni_api_test")
P4 = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("t", "<i4"), ("pad", "<i4")])
OUT = np.dtype([("i", "<i4"), ("j", "<i4"), ("h", "<f4")])


def _zunit(z, fp32=False):
    """(multiplier, LAS z scale) of the height unit nni::performNNI hands
    the engine for a work package whose heights read so far are z: the
    coarsest 10^-k (k <= 9) of which every height is a whole multiple, if
    the largest |height| fits zlim units; else 2^e metres, e the smallest
    with every |height| within zlim units.  zlim: 2^23 for FP32 two-sum,
    else 2^30."""
    zlim = 2.0 ** 23 if fp32 else 2.0 ** 30
    zmax = float(np.abs(z).max(initial=0.0))
    for k in range(10):
        u = z * 10.0 ** k
        if np.all(np.abs(u - np.rint(u)) <= 1e-6):
            if zmax * 10.0 ** k <= zlim:
                return 10.0 ** k, 1.0 / 10.0 ** k
            break
    e = -64
    while e < 64 and zmax > np.ldexp(zlim, e):
        e += 1
    return float(np.ldexp(1.0, -e)), float(np.ldexp(1.0, e))


def _run_api(tmp_path, pts, args):
    pin, pout = tmp_path / "pts.bin", tmp_path / "out.bin"
    pts.tofile(pin)
    cmd = [EXE, str(pin), str(pout)] + [f"{k}={v}" for k, v in args.items()]
    print("run", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.fromfile(pout, dtype=OUT)


@pytest.mark.parametrize("grid,n,cs,scale,slices,gox,tr", [
    (64, 3000, 0.5, 1, 1, 0, 0.0), (40, 2000, 2.0, 5, 3, 0, 0.0), (50, 4000, 1.0, 1, 2, 7, 0.0),
    (48, 6000, 0.5, 1, 4, 0, 1.0), (40, 4000, 1.0, 2, 3, 3, 2.0)])
def test_perform_nni_matches_oracle(tmp_path, grid, n, cs, scale, slices, gox, tr):
    """tr > 0: space-time slices (performNNI's timeRadius), one call per
    time t = 0..slices-1 on the same stream, as binRun's time loop."""
    tile = S.make_tile(40 + grid, n, grid, cell_size=cs, scale=scale, kind="clustered",
                       origin=(1000.0, 2000.0))
    p = tile.params
    x = tile.X * 0.01 + p.las_offset[0]
    y = tile.Y * 0.01 + p.las_offset[1]
    z = tile.Z * 0.01
    rng = np.random.default_rng(grid)
    t = np.sort(rng.integers(0, slices, tile.n)).astype(np.int32)  # time-sorted stream
    pts = np.zeros(tile.n, P4)
    pts["x"], pts["y"], pts["z"], pts["t"] = x, y, z, t
    total = grid + gox - 3  # the total grid clips the tile's last columns/rows
    args = dict(scale=scale, cell_size=cs, nrows=total, ncols=total, origin_x=p.origin_x,
                origin_y=p.origin_y, grid_origin_x=gox, grid_origin_y=gox, wp_nrows=grid,
                wp_ncols=grid, slices=slices, time_radius=tr)
    out = _run_api(tmp_path, pts, args)
    f = T.frame(p)
    upm = f["upm"]
    exp = []
    seen = np.zeros(tile.n, bool)
    for s in range(slices):
        sel = (t == s) if tr == 0 else np.ones(tile.n, bool)
        # the heights read so far: each call's slice, or every point within
        # 2r + 1 of curTime (space-time)
        seen |= (t == s) if tr == 0 else (np.abs(t - s) <= 2 * tr + 1)
        zmul, zs = _zunit(z[seen])
        X = np.rint((x[sel] - p.origin_x) * upm).astype(np.int32)
        Y = np.rint((y[sel] - p.origin_y) * upm).astype(np.int32)
        Z = np.rint(z[sel] * zmul).astype(np.int32)
        q = T.TileParams(nrows=grid, ncols=grid, cell_size=cs, scale=scale,
                         origin_x=p.origin_x, origin_y=p.origin_y,
                         las_scale=(1.0 / upm, 1.0 / upm, zs),
                         las_offset=(p.origin_x, p.origin_y, 0.0))
        if tr == 0:
            ref = O.run(q, X, Y, Z)["h_fixed"]
        else:
            ref = O.run_time(q, X, Y, Z, t, float(s), tr)["h_fixed"]
        for i in range(grid):          # column-major, as nni.cpp:836-860
            for j in range(grid):
                if i + gox < total and j + gox < total:
                    exp.append((i + gox, j + gox, ref[j, i]))
    exp = np.array(exp, dtype=OUT)
    assert out.shape == exp.shape
    assert np.array_equal(out["i"], exp["i"]) and np.array_equal(out["j"], exp["j"])
    assert np.array_equal(out["h"].view(np.uint32), exp["h"].view(np.uint32))


def test_delta_and_delta_squared_agree(tmp_path):
    """performNNI with useDelta (slices kept between calls) and without
    (Delta^2: every slice searched again) write the same cells, over six
    calls at time radius 2 on one stream."""
    grid, n = 40, 5000
    tile = S.make_tile(77, n, grid, cell_size=1.0, kind="clustered", origin=(500.0, 800.0))
    p = tile.params
    rng = np.random.default_rng(77)
    pts = np.zeros(tile.n, P4)
    pts["x"] = tile.X * 0.01 + p.las_offset[0]
    pts["y"] = tile.Y * 0.01 + p.las_offset[1]
    pts["z"] = tile.Z * 0.01
    pts["t"] = np.sort(rng.integers(0, 8, tile.n)).astype(np.int32)
    args = dict(cell_size=1.0, nrows=grid, ncols=grid, origin_x=p.origin_x, origin_y=p.origin_y,
                slices=6, time_radius=2.0)
    a = _run_api(tmp_path, pts, dict(args, delta=1))
    b = _run_api(tmp_path, pts, dict(args, delta=0))
    assert a.shape == (6 * grid * grid,)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert (a["h"] != T.NODATA).mean() > 0.5


@pytest.mark.parametrize("accum,zlo", [(0, 0.5), (1, 2.0)])
def test_submillimetre_heights_within_1e5(tmp_path, accum, zlo):
    """Point4D heights are doubles (nni.h:44-53): heights in [zlo, 50) m
    with sub-millimetre parts must come out within 1e-5 relative of the
    FP64 mean of the stolen sites' heights.  That mean is exact by
    linearity: the oracle is run on the millimetre part (LAS scale 1e-3)
    and on the remainder (scale 1e-9) with the same x/y, hence the same
    stolen pixels, and the two means are added.  Int64 sums (accum 0)
    carry heights to 2^-31 of the largest, FP32 two-sum (accum 1) to 2^-24;
    a unit of 1 mm would miss the bar by 10x at 5 m.  (Heights with at most
    9 decimals, as LAS integers give, are carried exactly.)"""
    grid, n, cs = 48, 4000, 0.5
    tile = S.make_tile(123, n, grid, cell_size=cs, kind="uniform", origin=(300.0, 700.0))
    p = tile.params
    rng = np.random.default_rng(5)
    x = tile.X * 0.01 + p.las_offset[0]
    y = tile.Y * 0.01 + p.las_offset[1]
    z = rng.uniform(zlo, 50.0, tile.n)
    pts = np.zeros(tile.n, P4)
    pts["x"], pts["y"], pts["z"] = x, y, z
    args = dict(scale=1, cell_size=cs, nrows=grid, ncols=grid, origin_x=p.origin_x,
                origin_y=p.origin_y, accum=accum)
    out = _run_api(tmp_path, pts, args)
    upm = T.frame(p)["upm"]
    X = np.rint((x - p.origin_x) * upm).astype(np.int32)
    Y = np.rint((y - p.origin_y) * upm).astype(np.int32)
    z_mm = np.rint(z * 1e3)
    z_rest = np.rint((z - z_mm * 1e-3) * 1e9)
    assert np.abs(z - (z_mm * 1e-3 + z_rest * 1e-9)).max() <= 5e-10

    def mean64(Zi, zs):
        q = T.TileParams(nrows=grid, ncols=grid, cell_size=cs, scale=1, origin_x=p.origin_x,
                         origin_y=p.origin_y, las_scale=(1.0 / upm, 1.0 / upm, zs),
                         las_offset=(p.origin_x, p.origin_y, 0.0))
        return O.run(q, X, Y, Zi.astype(np.int32))["h64"]

    a, b = mean64(z_mm, 1e-3), mean64(z_rest, 1e-9)
    valid = a != T.NODATA
    assert np.array_equal(valid, b != T.NODATA) and valid.mean() > 0.9
    ref = np.where(valid, a + b, T.NODATA)
    h = np.empty((grid, grid), np.float32)
    h[out["j"], out["i"]] = out["h"]
    assert np.array_equal(h != T.NODATA, valid)
    rel = np.abs(h[valid].astype(np.float64) - ref[valid]) / np.abs(ref[valid])
    print(f"accum {accum}: max rel {rel.max():.3e}, rms {np.sqrt(np.mean(rel ** 2)):.3e}")
    assert rel.max() <= 1e-5, rel.max()
