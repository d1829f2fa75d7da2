"""Pins of the CPU oracle (no GPU).

The reference has no tests, fixtures or CPU path, so the oracle is pinned by
closed forms of the reference's own formulas, by brute force, by a library
cross-check (scipy cKDTree) and by an independent scatter formulation of the
query pass, each chosen so that a plausible slip (an index off by one, a
dropped border, a transposed axis, a strict/non-strict test) fails one.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

import pyoracle as O
import terranni_amd as T
import terranni_amd.synth as S

NO_D2 = O.NO_D2


def P(**kw):
    base = dict(nrows=16, ncols=16, cell_size=0.5, scale=1, las_scale=(0.01, 0.01, 0.01),
                las_offset=(0.0, 0.0, 0.0))
    base.update(kw)
    return T.TileParams(**base)


# ---- pixel frame: nni.cpp:879-886 and the glOrtho of nni.cpp:656-660 ------

@pytest.mark.parametrize("cs,scale,n,po,W,ps_u,x0,Rs,Rq", [
    # 512 x 512 cells of 2 m at scale 5, qr = 4.99: pixel 0.4 m,
    # po = ceil((4.99 - 2 * 0.4) / 0.4) + 1 = 12, W = 2560 + 24,
    # pixel 0 centre = -1 - 12 * 0.4 + 0.2 = -5.6 m
    (2.0, 5, 512, 12, 2584, 40, -560, 353, 352),
    # 4096 x 4096 cells of 0.5 m at scale 1, qr = 1.24: po = ceil(2.48) + 1 = 4,
    # pixel 0 centre = -0.25 - 2 + 0.25 = -2 m
    (0.5, 1, 4096, 4, 4104, 50, -200, 353, 87),
    (0.5, 1, 256, 4, 264, 50, -200, 353, 87),
    (0.5, 1, 2048, 4, 2056, 50, -200, 353, 87),
])
def test_frame_closed_forms(cs, scale, n, po, W, ps_u, x0, Rs, Rq):
    p = P(nrows=n, ncols=n, cell_size=cs, scale=scale)
    f = O.frame(p)
    assert (f["po"], f["W"], f["H"], f["ps_u"], f["x0"], f["y0"]) == (po, W, W, ps_u, x0, x0)
    assert f["upm"] == 100 and f["cs_u"] == round(cs * 100)
    # site square half-side 5/sqrt(2) m, query half-side qr/sqrt(2) m, in cm
    assert (f["Rs"], f["Rq"]) == (Rs, Rq)
    assert f["spacing_factor"] == 1


def test_frame_pixel_centres_land_on_queries():
    # odd scale: the centre pixel of each cell's scale x scale block sits
    # exactly on the query (i*cs, j*cs)
    for scale in (1, 3, 5, 7):
        p = P(cell_size=2.0, scale=scale)
        f = O.frame(p)
        for i in (0, 1, 7, 15):
            k = f["po"] + i * scale + scale // 2
            assert f["x0"] + k * f["ps_u"] == i * f["cs_u"]


def test_frame_even_scale_and_thirds():
    f = O.frame(P(cell_size=1.0, scale=2))
    # pixel 0.5 m; centres at half pixels around the query
    assert f["ps_u"] == 50 and f["po"] == 5 and (f["x0"] - 25) % 50 == 0
    f = O.frame(P(cell_size=2.0, scale=3))  # 2/3 m pixels need 1/300 m units
    assert f["upm"] == 300 and f["ps_u"] == 200 and f["cs_u"] == 600


def test_frame_engine_matches_oracle_on_a_sweep():
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = P(nrows=int(rng.integers(1, 300)), ncols=int(rng.integers(1, 300)),
              cell_size=float(rng.choice([0.25, 0.5, 1.0, 1.5, 2.0, 3.0])),
              scale=int(rng.integers(1, 8)), site_radius=float(rng.choice([1.0, 2.5, 5.0, 10.0])),
              algorithm=int(rng.integers(0, 2)))
        try:
            fo = O.frame(p)
        except ValueError:
            with pytest.raises(T.TnniError):
                T.frame(p)
            continue
        fe = T.frame(p)
        assert fe == fo


def test_points_to_units_is_exact_for_las_integers():
    # LAS int32 at 0.01 m with a UTM-like offset; the origin on the same grid
    rng = np.random.default_rng(1)
    X = rng.integers(-10**6, 10**6, 5000).astype(np.int32)
    Y = rng.integers(-10**6, 10**6, 5000).astype(np.int32)
    p = P(nrows=8001, ncols=8001, cell_size=2.5, las_offset=(512345.67, 4123456.78, 0.0),
          origin_x=512345.67 - 10000.0, origin_y=4123456.78 - 10000.0)
    xu, yu, keep = O.points_to_units(p, X, Y)
    assert keep.all()
    assert np.array_equal(xu, X.astype(np.int64) + 10**6)
    assert np.array_equal(yu, Y.astype(np.int64) + 10**6)


# ---- labels ------------------------------------------------------------------

def _tiny(seed, n, grid, cs=0.5, scale=1, alg=0, kind="uniform"):
    t = S.make_tile(seed, n, grid, cell_size=cs, scale=scale, kind=kind)
    t.params.algorithm = alg
    return t


@pytest.mark.parametrize("seed,n,grid,cs,scale,alg", [
    (1, 200, 12, 0.5, 1, 0), (2, 300, 10, 2.0, 5, 0), (3, 150, 9, 1.0, 2, 1),
    (4, 400, 16, 0.5, 1, 1), (5, 50, 7, 2.0, 3, 0), (6, 1, 5, 0.5, 1, 0),
])
def test_labels_raster_equals_brute_force(seed, n, grid, cs, scale, alg):
    t = _tiny(seed, n, grid, cs, scale, alg)
    la, da = O.labels(t.params, t.X, t.Y)
    lb, db = O.labels(t.params, t.X, t.Y, brute=True)
    assert np.array_equal(la, lb) and np.array_equal(da, db)


@pytest.mark.parametrize("seed,alg", [(7, 0), (8, 1), (9, 0)])
def test_labels_match_scipy_ckdtree(seed, alg):
    """Library cross-check: the sites inside the region of each pixel
    centre by cKDTree.query_ball_point (p=inf for squares, p=2 for disks),
    then the smallest (d2, index)."""
    t = _tiny(seed, 600, 20, alg=alg, kind="clustered")
    f = O.frame(t.params)
    xu, yu, keep = O.points_to_units(t.params, t.X, t.Y)
    idx = np.nonzero(keep)[0]
    pts = np.stack([xu[idx], yu[idx]], 1).astype(np.float64)
    tree = cKDTree(pts)
    k, l = np.meshgrid(np.arange(f["W"]), np.arange(f["H"]))
    cen = np.stack([f["x0"] + k.ravel() * f["ps_u"], f["y0"] + l.ravel() * f["ps_u"]], 1)
    if alg == 0:
        hits = tree.query_ball_point(cen, r=f["Rs"], p=np.inf)
    else:
        hits = tree.query_ball_point(cen, r=math.sqrt(f["Rs2"]), p=2)
    lab_ref = np.full(cen.shape[0], -1, np.int64)
    d2_ref = np.full(cen.shape[0], NO_D2, np.int64)
    for q, h in enumerate(hits):
        if not h:
            continue
        h = np.asarray(h)
        d = ((pts[h] - cen[q]) ** 2).sum(1).astype(np.int64)
        if alg == 1:
            ok = d <= f["Rs2"]
            h, d = h[ok], d[ok]
            if h.size == 0:
                continue
        best = np.lexsort((idx[h], d))[0]
        lab_ref[q], d2_ref[q] = idx[h][best], d[best]
    lab, d2 = O.labels(t.params, t.X, t.Y)
    assert np.array_equal(lab.ravel(), lab_ref)
    assert np.array_equal(d2.ravel().astype(np.int64), d2_ref)


def _one_point_params(**kw):
    return P(nrows=8, ncols=8, origin_x=0.0, origin_y=0.0, **kw)


def test_single_site_labels_its_square():
    p = _one_point_params()
    f = O.frame(p)
    X = np.array([137], np.int32)  # 1.37 m
    Y = np.array([212], np.int32)
    lab, d2 = O.labels(p, X, Y)
    k, l = np.meshgrid(np.arange(f["W"]), np.arange(f["H"]))
    dx = f["x0"] + k * f["ps_u"] - 137
    dy = f["y0"] + l * f["ps_u"] - 212
    inside = (np.abs(dx) <= f["Rs"]) & (np.abs(dy) <= f["Rs"])
    assert np.array_equal(lab == 0, inside)
    assert np.array_equal(d2[inside], (dx * dx + dy * dy)[inside].astype(np.uint32))


def test_two_sites_split_on_the_bisector_lower_index_on_ties():
    p = _one_point_params()
    f = O.frame(p)
    # two sites symmetric about the vertical line x = 100 units (a pixel column)
    X = np.array([150, 50], np.int32)
    Y = np.array([100, 100], np.int32)
    lab, _ = O.labels(p, X, Y)
    k, l = np.meshgrid(np.arange(f["W"]), np.arange(f["H"]))
    px = f["x0"] + k * f["ps_u"]
    py = f["y0"] + l * f["ps_u"]
    cov0 = (np.abs(px - 150) <= f["Rs"]) & (np.abs(py - 100) <= f["Rs"])
    cov1 = (np.abs(px - 50) <= f["Rs"]) & (np.abs(py - 100) <= f["Rs"])
    both = cov0 & cov1
    assert (lab[both & (px > 100)] == 0).all()
    assert (lab[both & (px < 100)] == 1).all()
    assert (lab[both & (px == 100)] == 0).all()  # equidistant: lower index


def test_duplicates_and_empty():
    p = _one_point_params()
    X = np.array([200, 200, 200], np.int32)
    Y = np.array([300, 300, 300], np.int32)
    lab, _ = O.labels(p, X, Y)
    assert set(np.unique(lab)) <= {-1, 0}
    e = np.zeros(0, np.int32)
    lab, d2 = O.labels(p, e, e)
    assert (lab == -1).all() and (d2 == NO_D2).all()


def test_lattice_checkerboard():
    """Sites on a 1 m lattice: every pixel centre's label is the lattice
    point nearest to it (ties to the lower index = row-major order)."""
    p = P(nrows=10, ncols=10, cell_size=0.5, scale=1)
    f = O.frame(p)
    g = np.arange(-2, 8)
    xs, ys = np.meshgrid(g * 100, g * 100)
    X, Y = xs.ravel().astype(np.int32), ys.ravel().astype(np.int32)
    lab, _ = O.labels(p, X, Y)
    k, l = np.meshgrid(np.arange(f["W"]), np.arange(f["H"]))
    px = f["x0"] + k * f["ps_u"]
    py = f["y0"] + l * f["ps_u"]
    # nearest lattice coordinate, ties (half-way) to the lower one
    nx = np.ceil(px / 100.0 - 0.5).astype(int)
    ny = np.ceil(py / 100.0 - 0.5).astype(int)
    exp = (ny + 2) * 10 + (nx + 2)
    assert np.array_equal(lab, exp)


# ---- query pass + accumulation ---------------------------------------------

def scatter_reference(params, X, Y, Z):
    """Independent formulation of the query pass: pixel-major scatter.
    Each labelled pixel hands its site's height to every cell whose region
    covers it and which is at least as close to it as its site."""
    f = O.frame(params)
    lab, d2 = O.labels(params, X, Y, brute=True)
    nr, nc = params.nrows, params.ncols
    num = np.zeros((nr, nc), np.int64)
    cnt = np.zeros((nr, nc), np.int64)
    for l in range(f["H"]):
        for k in range(f["W"]):
            s = lab[l, k]
            if s < 0:
                continue
            px, py = f["x0"] + k * f["ps_u"], f["y0"] + l * f["ps_u"]
            for j in range(nr):
                dy = py - j * f["cs_u"]
                if abs(dy) > f["Rqb"]:
                    continue
                for i in range(nc):
                    dx = px - i * f["cs_u"]
                    if abs(dx) > f["Rqb"]:
                        continue
                    dq = dx * dx + dy * dy
                    if f["algorithm"] == 1 and dq > f["Rq2"]:
                        continue
                    if dq <= int(d2[l, k]):
                        num[j, i] += int(Z[s])
                        cnt[j, i] += 1
    return num, cnt


@pytest.mark.parametrize("seed,n,grid,cs,scale,alg", [
    (11, 60, 6, 0.5, 1, 0), (12, 80, 5, 2.0, 3, 0), (13, 40, 6, 1.0, 2, 1), (14, 100, 7, 0.5, 1, 1),
])
def test_heights_match_scatter_formulation(seed, n, grid, cs, scale, alg):
    t = _tiny(seed, n, grid, cs, scale, alg)
    num, cnt = scatter_reference(t.params, t.X, t.Y, t.Z)
    out = O.run(t.params, t.X, t.Y, t.Z)
    assert np.array_equal(out["count"].astype(np.int64), cnt)
    valid = cnt > 0
    h = np.where(valid, num / np.maximum(cnt, 1) * 0.01, -9999.0)
    assert np.allclose(out["h64"][valid], h[valid], rtol=1e-12, atol=1e-12)
    assert (out["h64"][~valid] == -9999.0).all()
    assert out["S"] == int(cnt.sum())


def test_constant_field_gives_the_constant():
    t = _tiny(15, 3000, 40, kind="clustered")
    Z = np.full_like(t.Z, 1234)
    out = O.run(t.params, t.X, t.Y, Z)
    valid = out["count"] > 0
    assert valid.any()
    assert (out["h_fixed"][valid] == np.float32(12.34)).all()
    assert np.allclose(out["h64"][valid], 12.34, rtol=1e-14)


def test_single_site_height_everywhere_it_reaches():
    p = _one_point_params()
    out = O.run(p, np.array([300], np.int32), np.array([300], np.int32), np.array([-4567], np.int32))
    valid = out["count"] > 0
    assert valid.sum() > 0
    assert (out["h_fixed"][valid] == np.float32(-45.67)).all()
    assert (out["h_fixed"][~valid] == -9999.0).all()


def test_cell_on_its_own_pixel_is_always_reached():
    # a query sitting on a labelled pixel centre steals it (dq = 0 <= d2)
    t = _tiny(16, 2000, 30)
    out = O.run(t.params, t.X, t.Y, t.Z)
    f = O.frame(t.params)
    ks = f["po"] + np.arange(t.params.ncols)  # scale 1: pixel po + i is on query i
    lab = out["labels"][np.ix_(ks, ks)]
    assert ((lab >= 0) <= (out["count"] > 0)).all()


# ---- the reference's integer path (S0) ----------------------------------------

def test_s0_site_value():
    # floor((z + zOffset + 1) * zMult), zOffset = -pointMin (nni.cpp:539,577)
    p = P(point_min=10, multiplier=1000)
    assert O.s0_value(p, 1250) == 3500          # (12.5 - 10 + 1) * 1000
    assert O.s0_value(p, 1000) == 1000          # z == pointMin -> 1 * zMult
    p = P(point_min=-3, multiplier=100)
    assert O.s0_value(p, -250) == 150           # (-2.5 + 3 + 1) * 100


def test_s0_heights_truncating_divide():
    # two sites of equal reach with values 3500 and 3501: the truncating
    # integer divide (cuda_nni.cu:500) gives 3500 where both are stolen
    p = P(nrows=4, ncols=4, point_min=10, multiplier=1000)
    X = np.array([50, 50], np.int32)
    Y = np.array([50, 50], np.int32)
    Z = np.array([1250, 1250], np.int32)
    out = O.run(p, X, Y, Z, s0=True)
    valid = out["count"] > 0
    # de-scale (nni.cpp:841-846): 3500 / 1000 - (-pointMin) - 1 = 12.5
    assert (out["h_s0"][valid] == np.float32(12.5)).all()
    assert (out["h_s0"][~valid] == -9999.0).all()


# ---- the ABI library, without a GPU -------------------------------------------

def _declared_symbols():
    import os
    import re
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    names = set()
    for hname in ("terranni.h",):
        src = open(os.path.join(here, "include", hname)).read()
        names |= set(re.findall(r"\b(tnni_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = T.lib()
    names = _declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_errors_without_gpu():
    with pytest.raises(T.TnniError) as e:
        T.frame(P(scale=0))
    assert e.value.status == -1
    with pytest.raises(T.TnniError) as e:
        T.frame(P(cell_size=0.123456))  # no exact fixed-point grid
    assert e.value.status == -2
    # a query offset of 1 km (70710 units) does not square into 32 bits
    with pytest.raises(T.TnniError) as e:
        T.frame(P(query_radius=1000.0))
    assert e.value.status == -2
    with pytest.raises(ValueError):
        O.frame(P(query_radius=1000.0))
    assert T.frame(P(query_radius=300.0))["Rqb"] == O.frame(P(query_radius=300.0))["Rqb"] == 21213
    assert T.footprint(P(), 1000) > 12000


def test_context_creation_reports_missing_device():
    import ctypes
    import os
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    st = T.lib().tnni_ctx_create(ctypes.byref(T.CtxOpts()), ctypes.byref(h))
    assert st == -3 and not h.value
