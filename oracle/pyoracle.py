"""ctypes wrapper of the CPU oracle (oracle/nni_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results or as the timed CPU baseline.  The
product library never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
NO_D2 = 0xFFFFFFFF


class OrcDesc(ctypes.Structure):
    _fields_ = [
        ("origin_x", ctypes.c_double), ("origin_y", ctypes.c_double),
        ("las_scale", ctypes.c_double * 3), ("las_offset", ctypes.c_double * 3),
        ("cell_size", ctypes.c_float), ("scale", ctypes.c_int),
        ("nrows", ctypes.c_int), ("ncols", ctypes.c_int),
        ("site_radius", ctypes.c_float), ("query_radius", ctypes.c_float),
        ("algorithm", ctypes.c_int), ("units_per_m", ctypes.c_int),
        ("pixel_offset", ctypes.c_int), ("point_min", ctypes.c_int),
        ("multiplier", ctypes.c_int),
    ]


class OrcFrame(ctypes.Structure):
    _fields_ = [
        ("upm", ctypes.c_int), ("ps_u", ctypes.c_int), ("cs_u", ctypes.c_int),
        ("po", ctypes.c_int), ("W", ctypes.c_int), ("H", ctypes.c_int),
        ("x0", ctypes.c_int), ("y0", ctypes.c_int), ("Rs", ctypes.c_int), ("Rq", ctypes.c_int),
        ("Rs2", ctypes.c_int64), ("Rq2", ctypes.c_int64),
        ("Rsb", ctypes.c_int), ("Rqb", ctypes.c_int),
        ("spacing_factor", ctypes.c_int), ("algorithm", ctypes.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    L.orc_frame_compute.argtypes = [ctypes.POINTER(OrcDesc), ctypes.POINTER(OrcFrame)]
    L.orc_frame_compute.restype = ctypes.c_int
    L.orc_points_to_units.argtypes = [ctypes.POINTER(OrcDesc), ctypes.POINTER(OrcFrame),
                                      P, P, ctypes.c_int64, P, P, P]
    L.orc_points_to_units.restype = None
    L.orc_labels_raster.argtypes = [ctypes.POINTER(OrcFrame), P, P, P, ctypes.c_int64, P, P,
                                    ctypes.c_int]
    L.orc_labels_raster.restype = None
    L.orc_labels_brute.argtypes = [ctypes.POINTER(OrcFrame), P, P, P, ctypes.c_int64, P, P]
    L.orc_labels_brute.restype = None
    L.orc_heights.argtypes = [ctypes.POINTER(OrcDesc), ctypes.POINTER(OrcFrame), P, P, P,
                              P, P, P, P, P, ctypes.c_int]
    L.orc_heights.restype = None
    L.orc_s0_value.argtypes = [ctypes.POINTER(OrcDesc), ctypes.c_int32]
    L.orc_s0_value.restype = ctypes.c_uint32
    L.orc_slice_sites.argtypes = [ctypes.POINTER(OrcFrame), P, P, ctypes.c_int64, ctypes.c_float,
                                  ctypes.c_float, P, P]
    L.orc_slice_sites.restype = None
    L.orc_labels_raster_t.argtypes = [ctypes.POINTER(OrcFrame), P, P, P, P, ctypes.c_int64, P, P,
                                      ctypes.c_int]
    L.orc_labels_raster_t.restype = None
    L.orc_labels_brute_t.argtypes = [ctypes.POINTER(OrcFrame), P, P, P, P, ctypes.c_int64, P, P]
    L.orc_labels_brute_t.restype = None
    L.orc_heights_accumulate.argtypes = [ctypes.POINTER(OrcDesc), ctypes.POINTER(OrcFrame), P, P,
                                         P, ctypes.c_int64, P, P, ctypes.c_int]
    L.orc_heights_accumulate.restype = None
    L.orc_heights_finish.argtypes = [ctypes.POINTER(OrcDesc), P, P, P, P, P]
    L.orc_heights_finish.restype = None
    _lib = L
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def desc_from(params) -> OrcDesc:
    """Build the oracle descriptor from a terranni_amd.TileParams."""
    d = OrcDesc()
    d.origin_x, d.origin_y = params.origin_x, params.origin_y
    for k in range(3):
        d.las_scale[k] = params.las_scale[k]
        d.las_offset[k] = params.las_offset[k]
    d.cell_size = params.cell_size
    d.scale = params.scale
    d.nrows, d.ncols = params.nrows, params.ncols
    d.site_radius = params.site_radius
    d.query_radius = params.qr()
    d.algorithm = params.algorithm
    d.units_per_m = params.units_per_m
    d.pixel_offset = params.pixel_offset
    d.point_min = params.point_min
    d.multiplier = params.multiplier
    return d


def frame(params) -> dict:
    d = desc_from(params)
    f = OrcFrame()
    rc = lib().orc_frame_compute(ctypes.byref(d), ctypes.byref(f))
    if rc != 0:
        raise ValueError(f"oracle: geometry not representable (code {rc})")
    return f.as_dict()


def _prep(params, X, Y):
    d = desc_from(params)
    f = OrcFrame()
    rc = lib().orc_frame_compute(ctypes.byref(d), ctypes.byref(f))
    if rc != 0:
        raise ValueError(f"oracle: geometry not representable (code {rc})")
    X = np.ascontiguousarray(X, dtype=np.int32)
    Y = np.ascontiguousarray(Y, dtype=np.int32)
    n = X.shape[0]
    xu = np.empty(n, np.int32)
    yu = np.empty(n, np.int32)
    keep = np.empty(n, np.uint8)
    lib().orc_points_to_units(ctypes.byref(d), ctypes.byref(f), _p(X), _p(Y), n, _p(xu), _p(yu),
                              _p(keep))
    return d, f, xu, yu, keep


def points_to_units(params, X, Y):
    _, _, xu, yu, keep = _prep(params, X, Y)
    return xu, yu, keep.astype(bool)


def labels(params, X, Y, brute: bool = False, nthreads: int = 0):
    """(label[H,W] int32, d2[H,W] uint32) of the site pass."""
    d, f, xu, yu, keep = _prep(params, X, Y)
    d2 = np.empty((f.H, f.W), np.uint32)
    lab = np.empty((f.H, f.W), np.int32)
    if brute:
        lib().orc_labels_brute(ctypes.byref(f), _p(xu), _p(yu), _p(keep), xu.shape[0], _p(d2),
                               _p(lab))
    else:
        lib().orc_labels_raster(ctypes.byref(f), _p(xu), _p(yu), _p(keep), xu.shape[0], _p(d2),
                                _p(lab), nthreads)
    return lab, d2


def heights(params, lab, d2, Z, nthreads: int = 0, s0: bool = False):
    """dict with h_fixed (float32), h64 (float64), count (uint32), K, S and
    (if s0) h_s0 (float32), all [nrows, ncols]."""
    d = desc_from(params)
    f = OrcFrame()
    if lib().orc_frame_compute(ctypes.byref(d), ctypes.byref(f)) != 0:
        raise ValueError("oracle: geometry not representable")
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    lab = np.ascontiguousarray(lab, dtype=np.int32)
    d2 = np.ascontiguousarray(d2, dtype=np.uint32)
    shp = (params.nrows, params.ncols)
    hf = np.empty(shp, np.float32)
    h64 = np.empty(shp, np.float64)
    hs0 = np.empty(shp, np.float32) if s0 else None
    cnt = np.empty(shp, np.uint32)
    pairs = np.zeros(2, np.uint64)
    lib().orc_heights(ctypes.byref(d), ctypes.byref(f), _p(d2), _p(lab), _p(Z), _p(hf), _p(h64),
                      _p(hs0), _p(cnt), _p(pairs), nthreads)
    out = {"h_fixed": hf, "h64": h64, "count": cnt, "K": int(pairs[0]), "S": int(pairs[1])}
    if s0:
        out["h_s0"] = hs0
    return out


def run(params, X, Y, Z, nthreads: int = 0, s0: bool = False, brute: bool = False):
    lab, d2 = labels(params, X, Y, brute=brute, nthreads=nthreads)
    out = heights(params, lab, d2, Z, nthreads=nthreads, s0=s0)
    out["labels"] = lab
    out["d2"] = d2
    return out


def s0_value(params, Z: int) -> int:
    d = desc_from(params)
    return int(lib().orc_s0_value(ctypes.byref(d), int(Z)))


# one orc_acc per cell: sumZ int64, sum64 double, numer0, denom0, cnt, pad u32
ACC = np.dtype([("sumZ", "<i8"), ("sum64", "<f8"), ("numer0", "<u4"), ("denom0", "<u4"),
                ("cnt", "<u4"), ("pad", "<u4")])


def time_slices(time_radius: float):
    """j of performNNI's slice loop: -(int)timeRadius .. timeRadius
    (nni.cpp:910)."""
    j = -int(time_radius)
    out = []
    while j <= time_radius:
        out.append(j)
        j += 1
    return out


def run_time(params, X, Y, Z, T, cur_time: float, time_radius: float, nthreads: int = 0,
             s0: bool = False, brute: bool = False, keep_slices: bool = False):
    """Space-time tile (nni.cpp:902-936): slices t = cur_time + j (float),
    sites of each filtered by time and at their squared time distance, the
    queries of slice j at (j * upm)^2, sums over all slices, one divide.
    Returns the dict of heights(); keep_slices adds per-slice (labels, d2)."""
    d, f, xu, yu, keep = _prep(params, X, Y)
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    T = np.ascontiguousarray(T, dtype=np.int32)
    n = xu.shape[0]
    acc = np.zeros(params.nrows * params.ncols, ACC)
    pairs = np.zeros(2, np.uint64)
    ks = np.empty(n, np.uint8)
    dt2 = np.empty(n, np.uint32)
    slices = []
    for j in time_slices(time_radius):
        t = np.float32(np.float32(cur_time) + np.float32(j))
        lib().orc_slice_sites(ctypes.byref(f), _p(T), _p(keep), n, float(t), float(time_radius),
                              _p(ks), _p(dt2))
        d2 = np.empty((f.H, f.W), np.uint32)
        lab = np.empty((f.H, f.W), np.int32)
        if brute:
            lib().orc_labels_brute_t(ctypes.byref(f), _p(xu), _p(yu), _p(ks), _p(dt2), n, _p(d2),
                                     _p(lab))
        else:
            lib().orc_labels_raster_t(ctypes.byref(f), _p(xu), _p(yu), _p(ks), _p(dt2), n, _p(d2),
                                      _p(lab), nthreads)
        jj2 = (j * f.upm) ** 2
        lib().orc_heights_accumulate(ctypes.byref(d), ctypes.byref(f), _p(d2), _p(lab), _p(Z), jj2,
                                     _p(acc), _p(pairs), nthreads)
        if keep_slices:
            slices.append((lab, d2))
    shp = (params.nrows, params.ncols)
    hf = np.empty(shp, np.float32)
    h64 = np.empty(shp, np.float64)
    hs0 = np.empty(shp, np.float32) if s0 else None
    cnt = np.empty(shp, np.uint32)
    lib().orc_heights_finish(ctypes.byref(d), _p(acc), _p(hf), _p(h64), _p(hs0), _p(cnt))
    out = {"h_fixed": hf, "h64": h64, "count": cnt, "K": int(pairs[0]), "S": int(pairs[1])}
    if s0:
        out["h_s0"] = hs0
    if keep_slices:
        out["slices"] = slices
    return out
