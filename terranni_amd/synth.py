"""Seeded synthetic LiDAR tiles for the benchmark and the parity tests.

The reference ships no data and its generator (data-generator.py) is
unseeded Python 2 text output, so the workloads are defined here:

  uniform   points uniform over the tile (+ a border), as in the 100k-point
            1000 x 1000 m workload;
  clustered Matern cluster process: parents uniform over land, children
            uniform in discs of radius U[5, 40] m, 'water' discs covering
            30 % of the area hold no returns, 1 % of the points are exact xy
            duplicates of others with a new height (ties).
Heights everywhere: sum of 4 sinusoids (amplitude 5-50 m, wavelength
50-500 m) + N(0, 0.1 m) noise.  Coordinates and heights are LAS int32 at
0.01 m with a UTM-like offset, so the engine's FP64 de-scale is exercised.

Everything is numpy with np.random.default_rng(seed): the same seed gives
the same tile on the same machine.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import TileParams

LAS_SCALE = (0.01, 0.01, 0.01)
UTM_OFFSET = (500000.0, 4200000.0, 0.0)


@dataclass
class Tile:
    params: TileParams
    X: np.ndarray
    Y: np.ndarray
    Z: np.ndarray
    kind: str = "uniform"
    seed: int = 0

    @property
    def n(self) -> int:
        return int(self.X.shape[0])

    @property
    def cells(self) -> int:
        return self.params.nrows * self.params.ncols


class HeightField:
    """z(x, y) = sum_k A_k sin(2 pi (x cos t_k + y sin t_k) / L_k + p_k)."""

    def __init__(self, rng: np.random.Generator):
        self.A = rng.uniform(5.0, 50.0, 4)
        self.L = rng.uniform(50.0, 500.0, 4)
        self.T = rng.uniform(0.0, 2 * np.pi, 4)
        self.P = rng.uniform(0.0, 2 * np.pi, 4)

    def __call__(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        z = np.zeros_like(x, dtype=np.float64)
        for k in range(4):
            z += self.A[k] * np.sin(2 * np.pi * (x * np.cos(self.T[k]) + y * np.sin(self.T[k]))
                                    / self.L[k] + self.P[k])
        return z


def _to_las(x, y, z, origin):
    """Tile-local metres -> LAS int32 with offset UTM_OFFSET; the tile's
    origin sits at UTM_OFFSET + origin."""
    ox, oy = origin
    X = np.rint((x + ox) / LAS_SCALE[0]).astype(np.int64)
    Y = np.rint((y + oy) / LAS_SCALE[1]).astype(np.int64)
    Z = np.rint(z / LAS_SCALE[2]).astype(np.int64)
    assert np.abs(X).max(initial=0) < 2**31 and np.abs(Y).max(initial=0) < 2**31
    return X.astype(np.int32), Y.astype(np.int32), Z.astype(np.int32)


def _params(nrows, ncols, cell_size, scale, origin, **kw) -> TileParams:
    return TileParams(nrows=nrows, ncols=ncols, cell_size=cell_size, scale=scale,
                      origin_x=UTM_OFFSET[0] + origin[0], origin_y=UTM_OFFSET[1] + origin[1],
                      las_scale=LAS_SCALE, las_offset=UTM_OFFSET, **kw)


def _water_mask(rng, ext, frac, res=2.0):
    """Union of random discs covering `frac` of [0, ext)^2 on a res-metre raster."""
    g = max(1, int(np.ceil(ext / res)))
    mask = np.zeros((g, g), dtype=bool)
    if frac <= 0:
        return mask, res
    yy, xx = np.mgrid[0:g, 0:g]
    cx_all = (xx + 0.5) * res
    cy_all = (yy + 0.5) * res
    while mask.mean() < frac:
        r = rng.uniform(0.03, 0.12) * ext
        cx, cy = rng.uniform(0, ext, 2)
        mask |= (cx_all - cx) ** 2 + (cy_all - cy) ** 2 <= r * r
    return mask, res


def _in_water(mask, res, x, y):
    g = mask.shape[0]
    i = np.clip((x / res).astype(np.int64), 0, g - 1)
    j = np.clip((y / res).astype(np.int64), 0, g - 1)
    return mask[j, i]


def uniform_points(rng, n, lo, hi):
    x = rng.uniform(lo, hi, n)
    y = rng.uniform(lo, hi, n)
    return x, y


def clustered_points(rng, n, ext, n_parents, water=0.30, dup=0.01, border=0.0):
    """Matern cluster process over [-border, ext + border)^2 with water."""
    mask, res = _water_mask(rng, ext, water)
    if mask.all():
        # an extent of a few metres can fall wholly in one water disc: there
        # is no land to put parents on, so the water is dropped
        mask[:] = False
    lo, hi = -border, ext + border
    # parents uniform over land
    px = np.empty(0)
    py = np.empty(0)
    while px.size < n_parents:
        cx = rng.uniform(lo, hi, 2 * n_parents)
        cy = rng.uniform(lo, hi, 2 * n_parents)
        land = ~_in_water(mask, res, cx, cy)
        px = np.concatenate([px, cx[land]])
        py = np.concatenate([py, cy[land]])
    px, py = px[:n_parents], py[:n_parents]
    pr = rng.uniform(5.0, 40.0, n_parents)
    n_main = n - int(round(dup * n))
    xs, ys = [], []
    have = 0
    while have < n_main:
        m = int((n_main - have) * 1.5) + 16
        k = rng.integers(0, n_parents, m)
        rr = pr[k] * np.sqrt(rng.uniform(0, 1, m))
        th = rng.uniform(0, 2 * np.pi, m)
        x = px[k] + rr * np.cos(th)
        y = py[k] + rr * np.sin(th)
        ok = (x >= lo) & (x < hi) & (y >= lo) & (y < hi) & ~_in_water(mask, res, x, y)
        xs.append(x[ok])
        ys.append(y[ok])
        have += int(ok.sum())
    x = np.concatenate(xs)[:n_main]
    y = np.concatenate(ys)[:n_main]
    nd = n - n_main
    if nd > 0 and n_main > 0:
        src = rng.integers(0, n_main, nd)
        x = np.concatenate([x, x[src]])
        y = np.concatenate([y, y[src]])
    perm = rng.permutation(x.shape[0])
    return x[perm], y[perm]


def make_tile(seed: int, n: int, grid: int, cell_size: float = 0.5, scale: int = 1,
              kind: str = "uniform", origin=(0.0, 0.0), border: float | None = None,
              n_parents: int | None = None, **kw) -> Tile:
    """A square tile of grid x grid cells with n points (border points included)."""
    rng = np.random.default_rng(seed)
    ext = grid * cell_size
    if border is None:
        border = 5 * cell_size + 5.0  # the driver's work-package border, terranni.cpp:659
    field = HeightField(rng)
    if kind == "uniform":
        x, y = uniform_points(rng, n, -border, ext + border)
    elif kind == "clustered":
        if n_parents is None:
            n_parents = max(4, int(round(2000 * (ext / 2048.0) ** 2)))
        x, y = clustered_points(rng, n, ext, n_parents, border=border)
    else:
        raise ValueError(kind)
    z = field(x, y) + rng.normal(0.0, 0.1, x.shape[0])
    X, Y, Z = _to_las(x, y, z, origin)
    p = _params(grid, grid, cell_size, scale, origin, **kw)
    return Tile(p, X, Y, Z, kind=kind, seed=seed)


def workload_uniform_100k(seed: int = 0, n: int = 100_000, **kw) -> Tile:
    """100k points uniform in 1000 x 1000 m, 512 x 512 cells of 2 m, scale 5."""
    rng = np.random.default_rng(seed)
    field = HeightField(rng)
    x, y = uniform_points(rng, n, 0.0, 1000.0)
    z = field(x, y) + rng.normal(0.0, 0.1, n)
    X, Y, Z = _to_las(x, y, z, (0.0, 0.0))
    p = _params(512, 512, 2.0, 5, (0.0, 0.0), **kw)
    return Tile(p, X, Y, Z, kind="uniform", seed=seed)


def workload_single_4m(seed: int = 1, n: int = 4_000_000, grid: int = 4096, **kw) -> Tile:
    """4M clustered points (2k parents, 30 % water, 1 % duplicates),
    4096 x 4096 cells of 0.5 m at scale 1."""
    rng = np.random.default_rng(seed)
    ext = grid * 0.5
    field = HeightField(rng)
    n_parents = max(4, int(round(2000 * (grid / 4096.0) ** 2)))
    x, y = clustered_points(rng, n, ext, n_parents, border=0.0)
    z = field(x, y) + rng.normal(0.0, 0.1, x.shape[0])
    X, Y, Z = _to_las(x, y, z, (0.0, 0.0))
    p = _params(grid, grid, 0.5, 1, (0.0, 0.0), **kw)
    return Tile(p, X, Y, Z, kind="clustered", seed=seed)


def workload_batch_manifest(seed: int = 3, n_tiles: int = 256, median: float = 2e5,
                            sigma: float = 1.0, nmin: float = 1e4, nmax: float = 2e6,
                            grids=(256, 512, 1024, 2048)):
    """(seed, n, grid, kind) of each uneven tile of the batched workload."""
    rng = np.random.default_rng(seed)
    ns = np.clip(rng.lognormal(np.log(median), sigma, n_tiles), nmin, nmax).astype(np.int64)
    gs = rng.choice(np.asarray(grids), n_tiles)
    kinds = np.where(rng.uniform(0, 1, n_tiles) < 0.5, "uniform", "clustered")
    seeds = rng.integers(0, 2**31 - 1, n_tiles)
    return [dict(seed=int(s), n=int(n), grid=int(g), kind=str(k))
            for s, n, g, k in zip(seeds, ns, gs, kinds)]


def workload_batch(seed: int = 3, n_tiles: int = 256, scale_points: float = 1.0, **kw):
    """The batched uneven-tile workload: list of Tiles."""
    man = workload_batch_manifest(seed, n_tiles, **kw)
    tiles = []
    for k, m in enumerate(man):
        n = max(1, int(m["n"] * scale_points))
        tiles.append(make_tile(m["seed"], n, m["grid"], kind=m["kind"],
                               origin=(float(k % 16) * 1100.0, float(k // 16) * 1100.0)))
    return tiles


# ---- LAS files ---------------------------------------------------------------

# Point record lengths of the LAS point formats this writer produces; every
# format starts with int32 X, Y, Z and uint16 intensity.
LAS_RECORD_LENGTH = {0: 20, 1: 28, 2: 26, 3: 34, 6: 30, 7: 36, 8: 38}


def write_las(path, X, Y, Z, intensity=None, scale=LAS_SCALE, offset=(0.0, 0.0, 0.0),
              version=(1, 2), point_format=0):
    """Write a LAS 1.2 or 1.4 file (no VLRs) of raw int32 coordinates; the
    header's bounding box is that of the de-scaled points."""
    X = np.ascontiguousarray(X, np.int32)
    Y = np.ascontiguousarray(Y, np.int32)
    Z = np.ascontiguousarray(Z, np.int32)
    n = X.shape[0]
    I = np.zeros(n, np.uint16) if intensity is None else np.ascontiguousarray(intensity, np.uint16)
    rl = LAS_RECORD_LENGTH[point_format]
    hsize = 375 if version >= (1, 4) else 227
    h = bytearray(hsize)
    h[0:4] = b"LASF"
    h[24], h[25] = version
    h[26:58] = b"terranni_amd synth".ljust(32, b"\0")
    h[58:90] = b"terranni_amd".ljust(32, b"\0")
    np.frombuffer(memoryview(h)[94:96], np.uint16)[:] = hsize
    np.frombuffer(memoryview(h)[96:100], np.uint32)[:] = hsize
    h[104] = point_format
    np.frombuffer(memoryview(h)[105:107], np.uint16)[:] = rl
    legacy = n if (point_format <= 5 and n < 2**32) else 0
    np.frombuffer(memoryview(h)[107:111], np.uint32)[:] = legacy
    np.frombuffer(memoryview(h)[131:155], np.float64)[:] = scale
    np.frombuffer(memoryview(h)[155:179], np.float64)[:] = offset
    if n:
        xs = X.astype(np.float64) * scale[0] + offset[0]
        ys = Y.astype(np.float64) * scale[1] + offset[1]
        zs = Z.astype(np.float64) * scale[2] + offset[2]
        box = [xs.max(), xs.min(), ys.max(), ys.min(), zs.max(), zs.min()]
    else:
        box = [0.0] * 6
    np.frombuffer(memoryview(h)[179:227], np.float64)[:] = box
    if hsize >= 375:
        np.frombuffer(memoryview(h)[247:255], np.uint64)[:] = n
    rec = np.zeros((n, rl), np.uint8)
    rec[:, 0:4] = X.view(np.uint8).reshape(n, 4)
    rec[:, 4:8] = Y.view(np.uint8).reshape(n, 4)
    rec[:, 8:12] = Z.view(np.uint8).reshape(n, 4)
    rec[:, 12:14] = I.view(np.uint8).reshape(n, 2)
    with open(path, "wb") as f:
        f.write(bytes(h))
        f.write(rec.tobytes())


def las_scene(seed: int, n: int, width: float, height: float, origin=(500000.0, 4200000.0),
              n_files: int = 3, kind: str = "uniform"):
    """A small LiDAR survey split into n_files strips along x: LAS int32 at
    0.01 m with offset `origin`; returns [(X, Y, Z, intensity)] per strip in
    file order (the points of strip k have x in its x range)."""
    rng = np.random.default_rng(seed)
    field = HeightField(rng)
    if kind == "uniform":
        x = rng.uniform(0, width, n)
        y = rng.uniform(0, height, n)
    else:
        x, y = clustered_points(rng, n, max(width, height), max(4, n // 2000), border=0.0)
        keep = (x < width) & (y < height)
        x, y = x[keep], y[keep]
    z = field(x, y) + rng.normal(0.0, 0.1, x.shape[0])
    X = np.rint(x / LAS_SCALE[0]).astype(np.int32)
    Y = np.rint(y / LAS_SCALE[1]).astype(np.int32)
    Z = np.rint(z / LAS_SCALE[2]).astype(np.int32)
    edges = np.linspace(0, width, n_files + 1)
    out = []
    for k in range(n_files):
        m = (x >= edges[k]) & (x < edges[k + 1])
        out.append((X[m], Y[m], Z[m], np.zeros(int(m.sum()), np.uint16)))
    return out


# ---- streamed terrain (fBm) --------------------------------------------------

def _hash01(ix, iy, octave, seed):
    """Deterministic value in [0, 1) per lattice node (SplitMix64-style mix)."""
    h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
         ^ iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
         ^ np.uint64((octave * 0x165667B19E3779F9 + seed * 0x27D4EB2F165667C5) & 0xFFFFFFFFFFFFFFFF))
    h ^= h >> np.uint64(30)
    h *= np.uint64(0xBF58476D1CE4E5B9)
    h ^= h >> np.uint64(27)
    h *= np.uint64(0x94D049BB133111EB)
    h ^= h >> np.uint64(31)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def fbm_height(x, y, seed: int, hurst: float = 0.8, octaves: int = 12, base_m: float = 4096.0,
               amp_m: float = 300.0):
    """Fractional Brownian terrain evaluated point by point: octaves of
    seeded value noise (smoothstep-interpolated lattice values), octave o at
    wavelength base_m / 2^o and amplitude amp_m * 2^(-hurst * o)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.zeros_like(x)
    with np.errstate(over="ignore"):
        for o in range(octaves):
            f = (2.0 ** o) / base_m
            u, v = x * f, y * f
            ix, iy = np.floor(u).astype(np.int64), np.floor(v).astype(np.int64)
            fx, fy = u - ix, v - iy
            sx, sy = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy)
            a = _hash01(ix, iy, o, seed)
            b = _hash01(ix + 1, iy, o, seed)
            c = _hash01(ix, iy + 1, o, seed)
            d = _hash01(ix + 1, iy + 1, o, seed)
            n = a + (b - a) * sx + (c - a) * sy + (a - b - c + d) * sx * sy
            z += amp_m * 2.0 ** (-hurst * o) * (2.0 * n - 1.0)
    return z


def _terrain_cell(seed: int, r: int, c: int, n: int, tile_m: float):
    """The points of core cell (r, c) of the streamed terrain: n uniform
    points in its tile_m square and their height noise, from the cell's own
    seed, so that every tile that overlaps the cell regenerates the same
    points."""
    rng = np.random.default_rng([seed, r, c])
    x = rng.uniform(c * tile_m, (c + 1) * tile_m, n)
    y = rng.uniform(r * tile_m, (r + 1) * tile_m, n)
    return x, y, rng.normal(0.0, 0.05, n)


def _terrain_tile(job):
    """Write LAS tile (r, c): the points of its core cell and of the
    neighbouring cells' points within `margin` of it (the overlap, points
    written to both tiles as exact duplicates), heights from the fBm."""
    out_dir, seed, r, c, nt, n_cell, tile_m, margin, origin = job
    import os
    lo_x, hi_x = c * tile_m - margin, (c + 1) * tile_m + margin
    lo_y, hi_y = r * tile_m - margin, (r + 1) * tile_m + margin
    xs, ys, es = [], [], []
    for rr in (r - 1, r, r + 1):
        for cc in (c - 1, c, c + 1):
            if 0 <= rr < nt and 0 <= cc < nt:
                x, y, e = _terrain_cell(seed, rr, cc, n_cell, tile_m)
                m = (x >= lo_x) & (x < hi_x) & (y >= lo_y) & (y < hi_y)
                xs.append(x[m])
                ys.append(y[m])
                es.append(e[m])
    x, y = np.concatenate(xs), np.concatenate(ys)
    z = fbm_height(x, y, seed) + np.concatenate(es)
    X = np.rint(x / LAS_SCALE[0]).astype(np.int32)
    Y = np.rint(y / LAS_SCALE[1]).astype(np.int32)
    Z = np.rint(z / LAS_SCALE[2]).astype(np.int32)
    path = os.path.join(out_dir, f"tile_r{r:03d}_c{c:03d}.las")
    write_las(path, X, Y, Z, None, offset=(origin[0], origin[1], 0.0))
    return path, int(X.shape[0])


def streamed_terrain_las(out_dir: str, seed: int = 5, extent_m: float = 5000.0,
                         n_points: int = 20_000_000, tile_m: float = 1000.0,
                         overlap: float = 0.05, origin=(500000.0, 4200000.0), workers: int = 16,
                         progress=None):
    """The streamed-terrain workload: n_points uniform over an extent_m
    square of fBm terrain (Hurst 0.8), cut into LAS tiles of tile_m with
    `overlap` of the tile side shared with each neighbour (points there are
    written to both files, exact duplicates).  Each tile_m core cell draws
    its n_points / cells points from its own seed, so the tiles are written
    independently, `workers` processes side by side.  Returns the list of
    files (sorted).  progress (optional): called with a line of text as
    tiles are written."""
    import multiprocessing as mp
    nt = int(np.ceil(extent_m / tile_m))
    n_cell = int(round(n_points / (nt * nt)))
    jobs = [(out_dir, seed, r, c, nt, n_cell, tile_m, overlap * tile_m, tuple(origin))
            for r in range(nt) for c in range(nt)]
    files, written = [], 0
    procs = max(1, min(workers, len(jobs), len(__import__("os").sched_getaffinity(0))))
    step = max(1, len(jobs) // 20)
    with mp.get_context("fork").Pool(procs) as pool:
        for k, (path, n) in enumerate(pool.imap_unordered(_terrain_tile, jobs, chunksize=1)):
            files.append(path)
            written += n
            if progress and ((k + 1) % step == 0 or k + 1 == len(jobs)):
                progress(f"terrain: {k + 1} of {len(jobs)} LAS tiles written ({written} points)")
    return sorted(files)
