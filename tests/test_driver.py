"""The streaming driver on the CPU: LAS 1.2 / 1.4 reading and binning into
work packages (terranni.cpp:198-368) through `terranni --bin-only`.

Every package must hold exactly the points within `border` of it
(border = 5 cells + site radius, terranni.cpp:656-657), in file order, then
sorted by time (stably); the plain engine's packages are nni::Point4D
streams in metres, the batched engine's hold the LAS integers in the first
file's scale / offset (a file in another frame is re-quantised to it)."""
import os
import subprocess

import numpy as np
import pytest

import terranni_amd.synth as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "terranni_amd", "_lib", "terranni")
P4 = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("t", "<i4"), ("pad", "<i4")])
RAW = np.dtype([("X", "<i4"), ("Y", "<i4"), ("Z", "<i4"), ("t", "<i4")])
OFFSET = (500000.0, 4200000.0, 0.0)
GRID = dict(cs=1.0, scale=2, sr=5.0, cols=250, rows=150, wpw=60, wph=40)

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")


def write_scene(d, seed=7, n=6000):
    """Three strips: LAS 1.2 format 0, LAS 1.2 format 3 at a 10x finer scale
    (re-quantised by the batched binning), LAS 1.4 format 6."""
    strips = S.las_scene(seed, n, 250.0, 150.0, origin=OFFSET[:2], n_files=3)
    specs = [((1, 2), 0, 1), ((1, 2), 3, 10), ((1, 4), 6, 1)]
    out = []
    for k, ((X, Y, Z, _), (ver, fmt, fine)) in enumerate(zip(strips, specs)):
        t = (np.arange(X.shape[0]) * 7 % 3).astype(np.uint16)
        sc = tuple(s / fine for s in S.LAS_SCALE)
        S.write_las(os.path.join(d, f"strip{k}.las"), X.astype(np.int64) * fine, Y.astype(np.int64) * fine,
                    Z.astype(np.int64) * fine, t, scale=sc, offset=OFFSET, version=ver, point_format=fmt)
        out.append((X, Y, Z, t, sc))
    return out


def run_driver(tmp_path, engine, extra=(), grid=None):
    las = tmp_path / "las"
    wp = tmp_path / f"wp_{engine}"
    out = tmp_path / f"out_{engine}"
    wp.mkdir(exist_ok=True)
    out.mkdir(exist_ok=True)
    g = grid or GRID
    cmd = [DRIVER, f"--tilepath={las}", f"--output={out}", f"--wp-path={wp}", f"--tmp-path={tmp_path}",
           f"--cell-size={g['cs']}", f"--scaling={g['scale']}", f"--site-radius={g['sr']}",
           f"--origin-x={OFFSET[0]}", f"--origin-y={OFFSET[1]}", f"--grid-cols={g['cols']}",
           f"--grid-rows={g['rows']}", f"--wp-width={g['wpw']}", f"--wp-height={g['wph']}",
           f"--engine={engine}", *extra]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return wp, out


def expected_packages(strips, grid=None):
    """Per package (row, col): the indices (file order) of the points within
    border of it, ordered by time then file order."""
    g = grid or GRID
    cs = float(np.float32(g["cs"]))
    border = float(np.float32(5 * np.float32(g["cs"]) + np.float32(g["sr"])))
    wpw, wph = int(g["wpw"] / cs) * cs, int(g["wph"] / cs) * cs
    cols = -(-g["cols"] // int(g["wpw"] / cs))
    rows = -(-g["rows"] // int(g["wph"] / cs))
    X = np.concatenate([s[0] for s in strips]).astype(np.float64)
    Y = np.concatenate([s[1] for s in strips]).astype(np.float64)
    t = np.concatenate([s[3] for s in strips]).astype(np.int64)
    x = X * S.LAS_SCALE[0] + OFFSET[0]
    y = Y * S.LAS_SCALE[1] + OFFSET[1]
    vx, vy = x - OFFSET[0], y - OFFSET[1]
    c0 = np.maximum(0, np.ceil((vx - border) / wpw - 1.0)).astype(int)
    c1 = np.minimum(cols - 1, np.floor((vx + border) / wpw)).astype(int)
    r0 = np.maximum(0, np.ceil((vy - border) / wph - 1.0)).astype(int)
    r1 = np.minimum(rows - 1, np.floor((vy + border) / wph)).astype(int)
    pk = {}
    for r in range(rows):
        for c in range(cols):
            idx = np.nonzero((c0 <= c) & (c <= c1) & (r0 <= r) & (r <= r1))[0]
            pk[(r, c)] = idx[np.argsort(t[idx], kind="stable")]
    return pk, rows, cols


def test_bin_plain_packages_hold_the_points_within_border(tmp_path):
    (tmp_path / "las").mkdir()
    strips = write_scene(str(tmp_path / "las"))
    wp, _ = run_driver(tmp_path, "plain", ["--bin-only"])
    pk, rows, cols = expected_packages(strips)
    x = np.concatenate([s[0].astype(np.float64) * S.LAS_SCALE[0] + OFFSET[0] for s in strips])
    z = np.concatenate([s[2].astype(np.float64) * S.LAS_SCALE[2] for s in strips])
    t = np.concatenate([s[3] for s in strips]).astype(np.int32)
    total = 0
    for (r, c), idx in pk.items():
        got = np.fromfile(wp / f"r{r}c{c}.tpie", dtype=P4)
        assert got.shape[0] == idx.shape[0], (r, c)
        # the fine-scale strip de-scales to the same metres up to rounding
        np.testing.assert_allclose(got["x"], x[idx], rtol=0, atol=1e-6)
        np.testing.assert_allclose(got["z"], z[idx], rtol=0, atol=1e-6)
        assert np.array_equal(got["t"], t[idx])
        total += got.shape[0]
    assert total > sum(s[0].shape[0] for s in strips)  # border copies exist


def test_bin_batched_packages_hold_las_integers_in_one_frame(tmp_path):
    (tmp_path / "las").mkdir()
    strips = write_scene(str(tmp_path / "las"))
    wp, _ = run_driver(tmp_path, "batched", ["--bin-only"])
    pk, rows, cols = expected_packages(strips)
    X = np.concatenate([s[0] for s in strips])
    Y = np.concatenate([s[1] for s in strips])
    Z = np.concatenate([s[2] for s in strips])
    t = np.concatenate([s[3] for s in strips]).astype(np.int32)
    for (r, c), idx in pk.items():
        got = np.fromfile(wp / f"r{r}c{c}.raw", dtype=RAW)
        assert np.array_equal(got["X"], X[idx]), (r, c)
        assert np.array_equal(got["Y"], Y[idx])
        assert np.array_equal(got["Z"], Z[idx])
        assert np.array_equal(got["t"], t[idx])
    meta = dict(line.split("=", 1) for line in (wp / "wp.meta").read_text().split())
    assert float(meta["scale0"]) == S.LAS_SCALE[0] and float(meta["offset0"]) == OFFSET[0]
    assert (int(meta["cols"]), int(meta["rows"])) == (cols, rows)


@pytest.mark.parametrize("block", [700, 4500, 0])
def test_bin_one_pass_over_many_packages(tmp_path, block):
    """1260 packages of 6 x 5 cells in one pass over the input, three files
    of about 2000 points: blocks of 700 points (each file read in several
    blocks, each block binned by all threads side by side), of 4500 (the
    first two files binned side by side, the third alone) and the default
    (all three side by side); every package holds exactly its points, in
    file order, then sorted by time."""
    (tmp_path / "las").mkdir()
    strips = write_scene(str(tmp_path / "las"))
    assert all(1500 < s[0].shape[0] < 2250 for s in strips)
    grid = dict(GRID, wpw=6, wph=5)
    extra = ["--bin-only"] + ([f"--bin-block={block}"] if block else [])
    wp, _ = run_driver(tmp_path, "batched", extra, grid=grid)
    pk, rows, cols = expected_packages(strips, grid)
    assert rows * cols == 1260
    X = np.concatenate([s[0] for s in strips])
    Z = np.concatenate([s[2] for s in strips])
    t = np.concatenate([s[3] for s in strips]).astype(np.int32)
    for (r, c), idx in pk.items():
        got = np.fromfile(wp / f"r{r}c{c}.raw", dtype=RAW)
        assert np.array_equal(got["X"], X[idx]), (r, c)
        assert np.array_equal(got["Z"], Z[idx]) and np.array_equal(got["t"], t[idx])


def test_2d_flag_zeroes_the_times(tmp_path):
    (tmp_path / "las").mkdir()
    write_scene(str(tmp_path / "las"))
    wp, _ = run_driver(tmp_path, "batched", ["--bin-only", "--2d"])
    for f in os.listdir(wp):
        if f.endswith(".raw"):
            assert not np.fromfile(wp / f, dtype=RAW)["t"].any()


def test_driver_rejects_bad_input(tmp_path):
    (tmp_path / "las").mkdir()
    (tmp_path / "las" / "bad.las").write_bytes(b"NOTLAS" + bytes(300))
    r = subprocess.run([DRIVER, f"--tilepath={tmp_path / 'las'}", f"--output={tmp_path}", "--bin-only"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "not a LAS file" in r.stderr
    r = subprocess.run([DRIVER, "--bogus=1"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "unknown option" in r.stderr
