"""ctypes bindings of the MI355X discrete-Sibson tile engine.

The engine itself is C++/HIP behind the C ABI declared in include/terranni.h
(built into terranni_amd/_lib/libterranni_hip.so).  This module loads that
library and wraps it for Python callers (tests, bench.py).  There is no
Python or CPU fallback: if the library is missing or no GPU is visible, the
calls raise.

The C++ API of the reference (nni.h: nni::init / performNNI / cleanup) is
provided by include/nni.h + terranni_amd/csrc/nni_shim.cpp on the same ABI.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libterranni_hip.so")

ALG_PLANES, ALG_CONES = 0, 1
ACC_FIXED64, ACC_FP32_2SUM, ACC_S0 = 0, 1, 2
FORM_AUTO, FORM_FUSED, FORM_SPLIT, FORM_SCATTER = 0, 1, 2, 3
NODATA = -9999.0
NO_D2 = 0xFFFFFFFF

STATUS = {
    0: "TNNI_OK", 1: "TNNI_PENDING", -1: "TNNI_E_INVALID", -2: "TNNI_E_GEOMETRY",
    -3: "TNNI_E_NODEVICE", -4: "TNNI_E_HIP", -5: "TNNI_E_NOMEM", -6: "TNNI_E_BUDGET",
    -7: "TNNI_E_SHUTDOWN",
}


class TnniError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class TileDesc(ctypes.Structure):
    """tnni_tile_desc of include/terranni.h."""
    _fields_ = [
        ("origin_x", ctypes.c_double), ("origin_y", ctypes.c_double),
        ("las_scale", ctypes.c_double * 3), ("las_offset", ctypes.c_double * 3),
        ("cell_size", ctypes.c_float), ("scale", ctypes.c_int),
        ("nrows", ctypes.c_int), ("ncols", ctypes.c_int),
        ("site_radius", ctypes.c_float), ("query_radius", ctypes.c_float),
        ("algorithm", ctypes.c_int), ("accum_mode", ctypes.c_int),
        ("units_per_m", ctypes.c_int), ("pixel_offset", ctypes.c_int),
        ("point_min", ctypes.c_int), ("multiplier", ctypes.c_int),
        ("bucket_px", ctypes.c_int), ("tile_cells", ctypes.c_int),
        ("form", ctypes.c_int),
    ]


class FrameInfo(ctypes.Structure):
    """tnni_frame_info of include/terranni.h."""
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


class CtxOpts(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int), ("n_slots", ctypes.c_int), ("hbm_budget", ctypes.c_uint64),
        ("max_batch_tiles", ctypes.c_int), ("batch_window_us", ctypes.c_int),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "completed", "batches", "max_inflight_tiles", "max_inflight_batches",
        "inflight_bytes", "budget_retries")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class TuneEntry(ctypes.Structure):
    _fields_ = [
        ("shape_class", ctypes.c_uint64), ("bucket_px", ctypes.c_int32),
        ("tile_cells", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32),
        ("candidates", ctypes.c_int32), ("ms_default", ctypes.c_float),
        ("ms_tuned", ctypes.c_float), ("form", ctypes.c_int32), ("pad", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


@dataclass
class TileParams:
    """One work package: nni::Settings + nni::WorkPackage + LAS scale/offset."""
    nrows: int
    ncols: int
    cell_size: float = 2.0
    scale: int = 5
    site_radius: float = 5.0
    query_radius: float | None = None   # None: 5*cs/2 - 0.01 (terranni.cpp:663)
    origin_x: float = 0.0
    origin_y: float = 0.0
    las_scale: tuple = (0.01, 0.01, 0.01)
    las_offset: tuple = (0.0, 0.0, 0.0)
    algorithm: int = ALG_PLANES
    accum_mode: int = ACC_FIXED64
    units_per_m: int = 0
    pixel_offset: int = -1
    point_min: int = 0
    multiplier: int = 1
    bucket_px: int = 0
    tile_cells: int = 0
    form: int = 0   # FORM_AUTO / FORM_FUSED / FORM_SPLIT / FORM_SCATTER (integer modes)

    def qr(self) -> float:
        if self.query_radius is not None:
            return float(self.query_radius)
        return 5 * float(np.float32(self.cell_size)) / 2.0 - 0.01

    def to_desc(self) -> TileDesc:
        d = TileDesc()
        d.origin_x, d.origin_y = self.origin_x, self.origin_y
        for k in range(3):
            d.las_scale[k] = self.las_scale[k]
            d.las_offset[k] = self.las_offset[k]
        d.cell_size = self.cell_size
        d.scale = self.scale
        d.nrows, d.ncols = self.nrows, self.ncols
        d.site_radius = self.site_radius
        d.query_radius = self.qr()
        d.algorithm = self.algorithm
        d.accum_mode = self.accum_mode
        d.units_per_m = self.units_per_m
        d.pixel_offset = self.pixel_offset
        d.point_min = self.point_min
        d.multiplier = self.multiplier
        d.bucket_px = self.bucket_px
        d.tile_cells = self.tile_cells
        d.form = self.form
        return d


_lib = None


def lib():
    """The loaded engine library.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.tnni_abi_version.restype = ctypes.c_int
    L.tnni_last_error.restype = ctypes.c_char_p
    L.tnni_last_error.argtypes = [P]
    L.tnni_frame.argtypes = [ctypes.POINTER(TileDesc), ctypes.POINTER(FrameInfo)]
    L.tnni_footprint.argtypes = [ctypes.POINTER(TileDesc), ctypes.c_uint64, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_uint64)]
    L.tnni_ctx_create.argtypes = [ctypes.POINTER(CtxOpts), ctypes.POINTER(P)]
    L.tnni_ctx_destroy.argtypes = [P]
    L.tnni_ctx_destroy.restype = None
    L.tnni_tile_run.argtypes = [P, ctypes.POINTER(TileDesc), P, P, P, ctypes.c_uint64, P, P, P]
    L.tnni_tile_run_dev.argtypes = [P, ctypes.POINTER(TileDesc), P, P, P, ctypes.c_uint64, P, P, P]
    L.tnni_tile_run_dev_noalloc.argtypes = [P, ctypes.POINTER(TileDesc), P, P, P, ctypes.c_uint64, P, P, P]
    L.tnni_batch_run_dev.argtypes = [P, ctypes.c_int, ctypes.POINTER(TileDesc), P, P, P,
                                     ctypes.POINTER(ctypes.c_uint64), P, P, P]
    L.tnni_dev_alloc.argtypes = [P, ctypes.c_uint64, ctypes.POINTER(P)]
    L.tnni_dev_free.argtypes = [P, P]
    L.tnni_h2d.argtypes = [P, P, P, ctypes.c_uint64]
    L.tnni_d2h.argtypes = [P, P, P, ctypes.c_uint64]
    L.tnni_d2d.argtypes = [P, P, P, ctypes.c_uint64]
    L.tnni_profile_enable.argtypes = [P, ctypes.c_int]
    L.tnni_profile_read.argtypes = [P, ctypes.POINTER(ctypes.c_uint64),
                                    ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.tnni_submit.argtypes = [P, ctypes.POINTER(TileDesc), P, P, P, ctypes.c_uint64, P, P, P,
                              ctypes.POINTER(ctypes.c_uint64)]
    L.tnni_submit_dev.argtypes = [P, ctypes.POINTER(TileDesc), P, P, P, ctypes.c_uint64, P, P, P,
                                  ctypes.POINTER(ctypes.c_uint64)]
    L.tnni_poll.argtypes = [P, ctypes.c_uint64]
    L.tnni_wait.argtypes = [P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    L.tnni_serve_stats_read.argtypes = [P, ctypes.POINTER(ServeStats)]
    L.tnni_serve_stats_reset.argtypes = [P]
    L.tnni_shape_class.argtypes = [ctypes.POINTER(TileDesc), ctypes.c_uint64, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_uint64)]
    L.tnni_autotune.argtypes = [P, ctypes.c_int, ctypes.POINTER(TileDesc), P, P, P,
                                ctypes.POINTER(ctypes.c_uint64), ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(TuneEntry), ctypes.POINTER(ctypes.c_int)]
    L.tnni_tune_read.argtypes = [P, ctypes.POINTER(TuneEntry), ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int)]
    L.tnni_tune_load.argtypes = [P, ctypes.POINTER(TuneEntry), ctypes.c_int]
    L.tnni_tune_clear.argtypes = [P]
    L.tnni_tile_run_time.argtypes = [P, ctypes.POINTER(TileDesc), P, P, P, P, ctypes.c_uint64,
                                     ctypes.c_float, ctypes.c_float, P]
    L.tnni_tile_run_time_dev.argtypes = [P, ctypes.POINTER(TileDesc), P, P, P, P, ctypes.c_uint64,
                                         ctypes.c_float, ctypes.c_float, P]
    L.tnni_time_cache_create.argtypes = [P, ctypes.c_int, ctypes.POINTER(P)]
    L.tnni_time_cache_destroy.argtypes = [P]
    L.tnni_time_cache_destroy.restype = None
    L.tnni_time_cache_clear.argtypes = [P]
    L.tnni_time_cache_stats.argtypes = [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    L.tnni_tile_run_time_cached.argtypes = [P, P, ctypes.POINTER(TileDesc), P, P, P, P, ctypes.c_uint64,
                                            ctypes.c_float, ctypes.c_float, P]
    L.tnni_tile_run_time_cached_dev.argtypes = [P, P, ctypes.POINTER(TileDesc), P, P, P, P,
                                                ctypes.c_uint64, ctypes.c_float, ctypes.c_float, P]
    for name in ("tnni_frame", "tnni_footprint", "tnni_ctx_create", "tnni_tile_run",
                 "tnni_tile_run_dev", "tnni_tile_run_dev_noalloc", "tnni_batch_run_dev", "tnni_dev_alloc", "tnni_dev_free",
                 "tnni_h2d", "tnni_d2h", "tnni_d2d", "tnni_profile_enable", "tnni_profile_read",
                 "tnni_submit", "tnni_submit_dev", "tnni_poll", "tnni_wait",
                 "tnni_serve_stats_read", "tnni_serve_stats_reset", "tnni_shape_class", "tnni_autotune",
                 "tnni_tune_read", "tnni_tune_load", "tnni_tune_clear", "tnni_tile_run_time",
                 "tnni_tile_run_time_dev", "tnni_time_cache_create", "tnni_time_cache_clear",
                 "tnni_time_cache_stats", "tnni_tile_run_time_cached", "tnni_tile_run_time_cached_dev"):
        getattr(L, name).restype = ctypes.c_int
    del i32p
    _lib = L
    return L


def _check(st: int, ctx=None):
    if st != 0:
        msg = lib().tnni_last_error(ctx)
        raise TnniError(st, msg.decode() if msg else "")


def shape_class(p: TileParams, n_points: int, batched: bool = True) -> int:
    """Shape class under which tuned sizes are kept (host only)."""
    c = ctypes.c_uint64()
    d = p.to_desc()
    _check(lib().tnni_shape_class(ctypes.byref(d), n_points, int(batched), ctypes.byref(c)))
    return c.value


def frame(p: TileParams) -> dict:
    """Pixel frame of a tile (host only, no GPU needed)."""
    f = FrameInfo()
    d = p.to_desc()
    _check(lib().tnni_frame(ctypes.byref(d), ctypes.byref(f)))
    return f.as_dict()


def footprint(p: TileParams, n_points: int, with_labels: bool = False) -> int:
    out = ctypes.c_uint64()
    d = p.to_desc()
    _check(lib().tnni_footprint(ctypes.byref(d), n_points, int(with_labels), ctypes.byref(out)))
    return out.value


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


@dataclass
class DevArray:
    ptr: int
    nbytes: int
    dtype: np.dtype
    shape: tuple
    ctx: "Context" = field(repr=False)

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        self.ctx._d2h(out, self)
        return out

    def free(self):
        if self.ptr:
            self.ctx._free(self.ptr)
            self.ptr = 0


class TimeCache:
    """Per point time, the 2-D Voronoi layer (site values and squared
    distances of the pixels) of a space-time tile, from which its slices are
    combined (tnni_time_cache; what the Delta render-buffer cache of
    nni.cpp:281-379 saves).  Close it before its context."""

    def __init__(self, h, ctx: "Context"):
        self.h = h
        self.ctx = ctx

    def clear(self):
        _check(lib().tnni_time_cache_clear(self.h))

    def stats(self):
        """(layers taken from the cache, layers searched)."""
        hits, misses = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().tnni_time_cache_stats(self.h, ctypes.byref(hits), ctypes.byref(misses)))
        return hits.value, misses.value

    def close(self):
        if self.h:
            lib().tnni_time_cache_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        # safe in either order with the context: a cache that outlives its
        # context holds nothing and only accepts destroy
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One engine context on one GPU (tnni_ctx)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, n_slots: int = 0,
                 max_batch_tiles: int = 0, batch_window_us: int = 0):
        o = CtxOpts(device, n_slots, hbm_budget, max_batch_tiles, batch_window_us)
        h = ctypes.c_void_p()
        _check(lib().tnni_ctx_create(ctypes.byref(o), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().tnni_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        _check(st, self.h)

    # ---- plain form on host buffers ----
    def tile_run(self, p: TileParams, X, Y, Z, labels: bool = False):
        """Run one tile; returns (heights[nrows, ncols], labels, d2) where
        labels/d2 are [H, W] rasters or None."""
        X = np.ascontiguousarray(X, dtype=np.int32)
        Y = np.ascontiguousarray(Y, dtype=np.int32)
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        n = X.shape[0]
        assert Y.shape[0] == n and Z.shape[0] == n
        fr = frame(p)
        h = np.empty((p.nrows, p.ncols), dtype=np.float32)
        lab = np.empty((fr["H"], fr["W"]), dtype=np.int32) if labels else None
        d2 = np.empty((fr["H"], fr["W"]), dtype=np.uint32) if labels else None
        d = p.to_desc()
        self._chk(lib().tnni_tile_run(self.h, ctypes.byref(d), _ptr(X), _ptr(Y), _ptr(Z), n,
                                      _ptr(h), _ptr(lab), _ptr(d2)))
        return h, lab, d2

    # ---- device memory ----
    def alloc(self, shape, dtype) -> DevArray:
        dtype = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dtype.itemsize
        p = ctypes.c_void_p()
        self._chk(lib().tnni_dev_alloc(self.h, nbytes, ctypes.byref(p)))
        return DevArray(p.value or 0, nbytes, dtype, tuple(shape), self)

    def upload(self, a: np.ndarray) -> DevArray:
        a = np.ascontiguousarray(a)
        d = self.alloc(a.shape, a.dtype)
        if a.nbytes:
            self._chk(lib().tnni_h2d(self.h, ctypes.c_void_p(d.ptr), _ptr(a), a.nbytes))
        return d

    def _d2h(self, out: np.ndarray, d: DevArray):
        if d.nbytes:
            self._chk(lib().tnni_d2h(self.h, _ptr(out), ctypes.c_void_p(d.ptr), d.nbytes))

    def copy(self, dst: DevArray, src: DevArray, nbytes: int | None = None):
        n = min(dst.nbytes, src.nbytes) if nbytes is None else nbytes
        self._chk(lib().tnni_d2d(self.h, ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), n))

    def _free(self, ptr):
        self._chk(lib().tnni_dev_free(self.h, ctypes.c_void_p(ptr)))

    def tile_run_dev(self, p, dX: DevArray, dY: DevArray, dZ: DevArray, n: int,
                     d_h: DevArray, d_lab: DevArray | None = None, d_d2: DevArray | None = None):
        """p: TileParams, or its TileDesc made once (p.to_desc()) for a
        caller that runs the same tile many times."""
        d = p if isinstance(p, TileDesc) else p.to_desc()
        vp = lambda a: ctypes.c_void_p(a.ptr) if a is not None else None  # noqa: E731
        self._chk(lib().tnni_tile_run_dev(self.h, ctypes.byref(d), vp(dX), vp(dY), vp(dZ), n,
                                          vp(d_h), vp(d_lab), vp(d_d2)))

    def tile_run_dev_noalloc(self, p: TileParams, dX: DevArray, dY: DevArray, dZ: DevArray, n: int,
                             d_h: DevArray, d_lab: DevArray | None = None, d_d2: DevArray | None = None):
        """tile_run_dev without any device allocation (fixed per-device
        scratch; tiles over its size raise TnniError with status -6)."""
        d = p.to_desc()
        vp = lambda a: ctypes.c_void_p(a.ptr) if a is not None else None  # noqa: E731
        self._chk(lib().tnni_tile_run_dev_noalloc(self.h, ctypes.byref(d), vp(dX), vp(dY), vp(dZ), n,
                                                  vp(d_h), vp(d_lab), vp(d_d2)))

    @staticmethod
    def batch_args(params, dXs, dYs, dZs, ns, d_hs, d_labs=None, d_d2s=None):
        """The C arguments of one tnni_batch_run_dev call, made once for a
        caller that runs the same batch many times (batch_run_args)."""
        k = len(params)
        descs = (TileDesc * k)(*[p.to_desc() for p in params])
        arr = lambda xs: (ctypes.c_void_p * k)(*[x.ptr if x is not None else None for x in xs])  # noqa
        nn = (ctypes.c_uint64 * k)(*ns)
        return (k, descs, arr(dXs), arr(dYs), arr(dZs), nn, arr(d_hs),
                arr(d_labs) if d_labs is not None else None,
                arr(d_d2s) if d_d2s is not None else None)

    def batch_run_args(self, args):
        self._chk(lib().tnni_batch_run_dev(self.h, *args))

    def batch_run_dev(self, params, dXs, dYs, dZs, ns, d_hs, d_labs=None, d_d2s=None):
        self.batch_run_args(self.batch_args(params, dXs, dYs, dZs, ns, d_hs, d_labs, d_d2s))

    # ---- space-time ----
    def tile_run_time(self, p: TileParams, X, Y, Z, T, cur_time: float, time_radius: float,
                      cache: "TimeCache" = None):
        """Heights of a space-time tile (host buffers): slices cur_time + j,
        j = -(int)r..r, sites within the time radius at their time distance.
        With a cache (and points sorted by time), the slices are combined
        from per-time layers, searching only the layers it lacks."""
        X = np.ascontiguousarray(X, dtype=np.int32)
        Y = np.ascontiguousarray(Y, dtype=np.int32)
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        T = np.ascontiguousarray(T, dtype=np.int32)
        h = np.empty((p.nrows, p.ncols), dtype=np.float32)
        d = p.to_desc()
        self._chk(lib().tnni_tile_run_time_cached(self.h, cache.h if cache else None, ctypes.byref(d),
                                                  _ptr(X), _ptr(Y), _ptr(Z), _ptr(T), X.shape[0],
                                                  cur_time, time_radius, _ptr(h)))
        return h

    def tile_run_time_dev(self, p: TileParams, dX, dY, dZ, dT, n: int, cur_time: float,
                          time_radius: float, d_h, cache: "TimeCache" = None):
        d = p.to_desc()
        self._chk(lib().tnni_tile_run_time_cached_dev(self.h, cache.h if cache else None, ctypes.byref(d),
                                                      dX.ptr, dY.ptr, dZ.ptr, dT.ptr, n, cur_time,
                                                      time_radius, d_h.ptr))

    def time_cache(self, slots: int = 0) -> "TimeCache":
        """A space-time cache on this context (tnni_time_cache); slots 0: the
        layers of one call."""
        h = ctypes.c_void_p()
        self._chk(lib().tnni_time_cache_create(self.h, slots, ctypes.byref(h)))
        return TimeCache(h, self)

    # ---- autotuning ----
    def autotune(self, params, dXs, dYs, dZs, ns, batched: bool = True, reps: int = 3):
        """Tune bucket / cell-tile sizes and workgroups per CU per shape
        class on device-resident sample tiles; returns one dict per class."""
        k = len(params)
        descs = (TileDesc * k)(*[p.to_desc() for p in params])
        arr = lambda xs: (ctypes.c_void_p * k)(*[x.ptr for x in xs])  # noqa
        nn = (ctypes.c_uint64 * k)(*ns)
        out = (TuneEntry * max(k, 1))()
        nc = ctypes.c_int()
        self._chk(lib().tnni_autotune(self.h, k, descs, arr(dXs), arr(dYs), arr(dZs), nn,
                                      int(batched), reps, out, ctypes.byref(nc)))
        return [out[i].as_dict() for i in range(nc.value)]

    def tune_table(self):
        n = ctypes.c_int()
        self._chk(lib().tnni_tune_read(self.h, None, 0, ctypes.byref(n)))
        out = (TuneEntry * max(n.value, 1))()
        self._chk(lib().tnni_tune_read(self.h, out, n.value, ctypes.byref(n)))
        return [out[i].as_dict() for i in range(n.value)]

    def tune_load(self, entries):
        arr = (TuneEntry * max(len(entries), 1))()
        for i, e in enumerate(entries):
            for f, _ in TuneEntry._fields_:
                if f in e:
                    setattr(arr[i], f, e[f])
        self._chk(lib().tnni_tune_load(self.h, arr, len(entries)))

    def tune_clear(self):
        self._chk(lib().tnni_tune_clear(self.h))

    # ---- served form ----
    def submit(self, p: TileParams, X, Y, Z, labels: bool = False, out: np.ndarray | None = None):
        """Queue a tile; returns a Pending whose arrays are filled once
        wait() returns.  The input arrays are kept alive until then.  out
        (optional): a float32 array of at least nrows*ncols elements that
        receives the heights (row-major), e.g. one of a caller's ring of
        reused buffers."""
        X = np.ascontiguousarray(X, dtype=np.int32)
        Y = np.ascontiguousarray(Y, dtype=np.int32)
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        fr = frame(p)
        if out is not None:
            if out.dtype != np.float32 or not out.flags.c_contiguous or out.size < p.nrows * p.ncols:
                raise ValueError("out must be a contiguous float32 array of at least nrows*ncols")
            h = out.reshape(-1)[:p.nrows * p.ncols].reshape(p.nrows, p.ncols)
        else:
            h = np.empty((p.nrows, p.ncols), dtype=np.float32)
        lab = np.empty((fr["H"], fr["W"]), dtype=np.int32) if labels else None
        d2 = np.empty((fr["H"], fr["W"]), dtype=np.uint32) if labels else None
        d = p.to_desc()
        t = ctypes.c_uint64()
        self._chk(lib().tnni_submit(self.h, ctypes.byref(d), _ptr(X), _ptr(Y), _ptr(Z), X.shape[0],
                                    _ptr(h), _ptr(lab), _ptr(d2), ctypes.byref(t)))
        return Pending(self, t.value, (X, Y, Z), h, lab, d2)

    def submit_dev(self, p: TileParams, dX: "DevArray", dY: "DevArray", dZ: "DevArray", n: int,
                   d_h: "DevArray", d_lab: "DevArray | None" = None,
                   d_d2: "DevArray | None" = None):
        """Queue a tile whose inputs and outputs live in device memory."""
        d = p.to_desc()
        t = ctypes.c_uint64()
        vp = lambda a: ctypes.c_void_p(a.ptr) if a is not None else None  # noqa: E731
        self._chk(lib().tnni_submit_dev(self.h, ctypes.byref(d), vp(dX), vp(dY), vp(dZ), n,
                                        vp(d_h), vp(d_lab), vp(d_d2), ctypes.byref(t)))
        return Pending(self, t.value, (dX, dY, dZ), None, None, None)

    def serve_stats(self) -> dict:
        s = ServeStats()
        self._chk(lib().tnni_serve_stats_read(self.h, ctypes.byref(s)))
        return s.as_dict()

    def serve_stats_reset(self):
        """Restart the in-flight high-water marks."""
        self._chk(lib().tnni_serve_stats_reset(self.h))

    # ---- profiling ----
    def profile(self, enable: bool):
        self._chk(lib().tnni_profile_enable(self.h, int(enable)))

    def profile_read(self, reset: bool = True):
        n = ctypes.c_uint64()
        ms = ctypes.c_double()
        self._chk(lib().tnni_profile_read(self.h, ctypes.byref(n), ctypes.byref(ms), int(reset)))
        return n.value, ms.value


@dataclass
class Pending:
    """A submitted tile (tnni_submit ticket)."""
    ctx: Context
    ticket: int
    inputs: tuple
    h: np.ndarray
    labels: np.ndarray | None
    d2: np.ndarray | None
    latency_s: float | None = None

    def done(self) -> bool:
        st = lib().tnni_poll(self.ctx.h, self.ticket)
        if st == 1:
            return False
        if st != 0:
            self.ctx._chk(st)
        return True

    def wait(self):
        lat = ctypes.c_double()
        st = lib().tnni_wait(self.ctx.h, self.ticket, ctypes.byref(lat))
        self.latency_s = lat.value
        self.inputs = ()
        self.ctx._chk(st)
        return self.h, self.labels, self.d2
