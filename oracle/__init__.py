"""Python view of the CPU oracle (test infrastructure only).

The oracle restates the reference's CPU scorer (src/sam_cpu.cpp,
src/detectors_cpu.cpp, src/detector.cpp, src/linalg.cpp, src/types.cpp) in C;
this module loads ``oracle/liboracle.so`` through ctypes and exposes numpy
functions. Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import it. The product path never does.

Cubes are BSQ float32 arrays of shape (bands, height, width) or
(bands, plane_stride); outputs are per pixel in raster order.
"""
from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libhsiref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_sz = C.c_size_t

DETECTION_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("target", "<i4"), ("score", "<f4")])

STATS_OK, STATS_EMPTY, STATS_BAND_MISMATCH, STATS_COV_NOT_PD, STATS_CORR_NOT_PD = range(5)


class _Ptr:
    """Accept None for a nullable pointer argument."""

    def __init__(self, base):
        self.base = base

    def from_param(self, obj):
        if obj is None:
            return None
        return self.base.from_param(obj)


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `make -C oracle` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_library_norms": (None, [_f32p, C.c_int, C.c_int, _f32p]),
        "orc_sam_angle": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "orc_sam_all": (None, [_f32p, _sz, _sz, _sz, C.c_int, _f32p, _f32p, C.c_int, _sz, _f32p]),
        "orc_sam_best": (None, [_f32p, _sz, _sz, _sz, C.c_int, _f32p, _f32p, C.c_int, _f32p, _i32p]),
        "orc_detect": (C.c_int64, [_f32p, _Ptr(_i32p), C.c_int, C.c_int, C.c_float, C.c_int,
                                   C.c_int, C.c_void_p, C.c_int64]),
        "orc_mean_spectrum": (C.c_int, [_f32p, _sz, C.c_int, _u8p, _f32p]),
        "orc_moments_accumulate": (None, [_f32p, _sz, _sz, _sz, C.c_int, _f64p, _f64p, _f64p]),
        "orc_moments_from_shifted": (None, [_f64p, _f64p, C.c_double, _f64p, C.c_int, _f64p, _f64p]),
        "orc_moments_mean": (None, [C.c_double, _f64p, C.c_int, _f64p]),
        "orc_moments_correlation": (None, [C.c_double, _f64p, C.c_int, C.c_double, _f64p]),
        "orc_moments_covariance": (None, [C.c_double, _f64p, _f64p, C.c_int, C.c_double, _f64p]),
        "orc_cholesky": (C.c_int, [_f64p, C.c_int]),
        "orc_cholesky_solve": (None, [_f64p, C.c_int, _f64p, _f64p]),
        "orc_cholesky_inverse": (None, [_f64p, C.c_int, _f64p]),
        "orc_triangular_inverse_lower": (None, [_f64p, C.c_int, _f64p]),
        "orc_quadratic_form": (C.c_double, [_f64p, _f64p, C.c_int]),
        "orc_build_statistics": (C.c_int, [C.c_double, _f64p, _f64p, C.c_int, _Ptr(_f32p), C.c_int,
                                           C.c_int, C.c_double, _f32p, _f32p, _f32p, _f32p, _f32p,
                                           C.c_char_p, C.c_int]),
        "orc_rx": (None, [_f32p, _sz, _sz, _sz, C.c_int, _f32p, _f32p, _f32p]),
        "orc_ace_best": (None, [_f32p, _sz, _sz, _sz, C.c_int, _f32p, _f32p, _f32p, _f32p, C.c_int,
                                _f32p, _Ptr(_i32p), _Ptr(_f32p)]),
        "orc_cem_best": (None, [_f32p, _sz, _sz, _sz, C.c_int, _f32p, C.c_int, _f32p, _Ptr(_i32p)]),
        "orc_gamma_p": (C.c_double, [C.c_double, C.c_double]),
        "orc_rx_threshold": (C.c_double, [C.c_int, C.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _flat_cube(cube):
    """(B, H, W) or (B, P) float32 -> (contiguous (B, stride) array, bands, stride)."""
    cube = np.ascontiguousarray(cube, dtype=np.float32)
    bands = cube.shape[0]
    flat = cube.reshape(bands, -1)
    return flat, bands, flat.shape[1]


def _ranges(pixels, threads):
    threads = max(1, min(int(threads), pixels if pixels else 1))
    step = (pixels + threads - 1) // threads
    return [(s, min(pixels, s + step)) for s in range(0, pixels, step)] or [(0, 0)]


def _run(fn, pixels, threads):
    """Call fn(p0, p1) over pixel ranges; ctypes drops the GIL during calls."""
    parts = _ranges(pixels, threads)
    if len(parts) == 1:
        fn(*parts[0])
        return
    with ThreadPoolExecutor(len(parts)) as pool:
        list(pool.map(lambda r: fn(*r), parts))


# ---- library / SAM ---------------------------------------------------------

def library_norms(targets):
    targets = np.ascontiguousarray(targets, dtype=np.float32)
    t, b = targets.shape
    out = np.empty(t, np.float32)
    lib().orc_library_norms(targets, t, b, out)
    return out


def sam_angle(dot, pixel_sq, norm):
    return lib().orc_sam_angle(dot, pixel_sq, norm)


def sam_best(cube, targets, norms=None, pixels=None, threads=1):
    flat, bands, stride = _flat_cube(cube)
    pixels = stride if pixels is None else pixels
    targets = np.ascontiguousarray(targets, dtype=np.float32)
    norms = library_norms(targets) if norms is None else np.ascontiguousarray(norms, np.float32)
    score = np.empty(stride, np.float32)
    target = np.empty(stride, np.int32)
    _run(lambda a, b: lib().orc_sam_best(flat, stride, a, b, bands, targets, norms,
                                         targets.shape[0], score, target), pixels, threads)
    return score[:pixels], target[:pixels]


def sam_all(cube, targets, norms=None):
    flat, bands, stride = _flat_cube(cube)
    targets = np.ascontiguousarray(targets, dtype=np.float32)
    norms = library_norms(targets) if norms is None else np.ascontiguousarray(norms, np.float32)
    out = np.empty((targets.shape[0], stride), np.float32)
    lib().orc_sam_all(flat, stride, 0, stride, bands, targets, norms, targets.shape[0], stride, out)
    return out


# ---- detect ----------------------------------------------------------------

def detect(score, target, height, width, threshold, nms_radius, higher_is_better,
           capacity=None):
    """Raster-order detections as a structured array, plus the found count."""
    score = np.ascontiguousarray(score, dtype=np.float32).reshape(-1)
    if target is not None:
        target = np.ascontiguousarray(target, dtype=np.int32).reshape(-1)
    found = lib().orc_detect(score, target, height, width, threshold, nms_radius,
                             int(bool(higher_is_better)), None, 0)
    cap = found if capacity is None else min(found, capacity)
    out = np.zeros(max(cap, 1), DETECTION_DTYPE)
    lib().orc_detect(score, target, height, width, threshold, nms_radius,
                     int(bool(higher_is_better)), out.ctypes.data, cap)
    return out[:cap], int(found)


def mean_spectrum(cube, mask):
    flat, bands, pixels = _flat_cube(cube)
    mask = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
    out = np.empty(bands, np.float32)
    if lib().orc_mean_spectrum(flat, pixels, bands, mask, out) != 0:
        raise ValueError("mean_spectrum: mask selected no pixels")
    return out


# ---- moments / linear algebra -------------------------------------------------

def moments(cube, pixels=None, threads=1):
    """(count, sum[B], sum_outer[B, B] lower triangle) as SpectralMoments::add leaves them."""
    flat, bands, stride = _flat_cube(cube)
    pixels = stride if pixels is None else pixels
    if threads <= 1:
        count = np.zeros(1, np.float64)
        s = np.zeros(bands, np.float64)
        so = np.zeros((bands, bands), np.float64)
        lib().orc_moments_accumulate(flat, stride, 0, pixels, bands, count, s, so)
        return float(count[0]), s, so
    parts = _ranges(pixels, threads)
    acc = [(np.zeros(1), np.zeros(bands), np.zeros((bands, bands))) for _ in parts]

    def work(i):
        a, b = parts[i]
        lib().orc_moments_accumulate(flat, stride, a, b, bands, *acc[i])

    with ThreadPoolExecutor(len(parts)) as pool:
        list(pool.map(work, range(len(parts))))
    # merged per SpectralMoments::merge (src/linalg.cpp:31-43); not bit-equal
    # to one serial pass, used only for timing
    return (float(sum(a[0][0] for a in acc)), sum(a[1] for a in acc), sum(a[2] for a in acc))


def moments_from_shifted(s1, s2_lower, count, reference):
    s1 = np.ascontiguousarray(s1, np.float64)
    s2 = np.ascontiguousarray(s2_lower, np.float64)
    ref = np.ascontiguousarray(reference, np.float64)
    b = s1.shape[0]
    out_s = np.empty(b, np.float64)
    out_o = np.empty((b, b), np.float64)
    lib().orc_moments_from_shifted(s1, s2, count, ref, b, out_s, out_o)
    return out_s, out_o


def covariance(count, s, so, loading=0.0):
    b = s.shape[0]
    out = np.empty((b, b), np.float64)
    lib().orc_moments_covariance(count, np.ascontiguousarray(s), np.ascontiguousarray(so), b,
                                 loading, out)
    return out


def correlation(count, so, loading=0.0):
    b = so.shape[0]
    out = np.empty((b, b), np.float64)
    lib().orc_moments_correlation(count, np.ascontiguousarray(so), b, loading, out)
    return out


def cholesky(a):
    a = np.array(a, dtype=np.float64, order="C", copy=True)
    ok = lib().orc_cholesky(a, a.shape[0])
    return bool(ok), a


def cholesky_solve(l, b):
    l = np.ascontiguousarray(l, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    lib().orc_cholesky_solve(l, l.shape[0], b, x)
    return x


def cholesky_inverse(l):
    l = np.ascontiguousarray(l, np.float64)
    inv = np.empty_like(l)
    lib().orc_cholesky_inverse(l, l.shape[0], inv)
    return inv


def triangular_inverse_lower(l):
    l = np.ascontiguousarray(l, np.float64)
    inv = np.empty_like(l)
    lib().orc_triangular_inverse_lower(l, l.shape[0], inv)
    return inv


def quadratic_form(m, y):
    m = np.ascontiguousarray(m, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().orc_quadratic_form(m, y, y.shape[0])


class Statistics:
    """DetectorStatistics (include/hsi/detector.hpp:62-85) as numpy arrays."""

    def __init__(self, bands, num_targets):
        self.bands = bands
        self.num_targets = num_targets
        self.mean = np.zeros(bands, np.float32)
        self.whitener = np.zeros((bands, bands), np.float32)
        self.target_white = np.zeros((max(num_targets, 1), bands), np.float32)
        self.target_norm2 = np.zeros(max(num_targets, 1), np.float32)
        self.cem_weight = np.zeros((max(num_targets, 1), bands), np.float32)
        self.status = STATS_EMPTY
        self.error = ""

    @property
    def valid(self):
        return self.status == STATS_OK


def build_statistics(count, s, so, targets, loading):
    """build_statistics (src/detector.cpp:47-152) from accumulated moments."""
    bands = s.shape[0]
    if targets is None or len(targets) == 0:
        tg, t, lb = None, 0, bands
    else:
        tg = np.ascontiguousarray(targets, np.float32)
        t, lb = tg.shape
    st = Statistics(bands, t)
    err = C.create_string_buffer(512)
    st.status = lib().orc_build_statistics(
        count, np.ascontiguousarray(s, np.float64), np.ascontiguousarray(so, np.float64), bands,
        tg, t, lb, loading, st.mean, st.whitener, st.target_white, st.target_norm2,
        st.cem_weight, err, 512)
    st.error = err.value.decode()
    return st


def statistics_of_cube(cube, targets, loading, pixels=None):
    count, s, so = moments(cube, pixels=pixels)
    return build_statistics(count, s, so, targets, loading)


# ---- second-order detectors ----------------------------------------------------

def rx(cube, stats, pixels=None, threads=1):
    flat, bands, stride = _flat_cube(cube)
    pixels = stride if pixels is None else pixels
    out = np.empty(stride, np.float32)
    _run(lambda a, b: lib().orc_rx(flat, stride, a, b, bands, stats.mean, stats.whitener, out),
         pixels, threads)
    return out[:pixels]


def ace_best(cube, stats, pixels=None, threads=1):
    flat, bands, stride = _flat_cube(cube)
    pixels = stride if pixels is None else pixels
    score = np.empty(stride, np.float32)
    target = np.empty(stride, np.int32)
    rxo = np.empty(stride, np.float32)
    _run(lambda a, b: lib().orc_ace_best(flat, stride, a, b, bands, stats.mean, stats.whitener,
                                         stats.target_white, stats.target_norm2,
                                         stats.num_targets, score, target, rxo), pixels, threads)
    return score[:pixels], target[:pixels], rxo[:pixels]


def cem_best(cube, stats, pixels=None, threads=1):
    flat, bands, stride = _flat_cube(cube)
    pixels = stride if pixels is None else pixels
    score = np.empty(stride, np.float32)
    target = np.empty(stride, np.int32)
    _run(lambda a, b: lib().orc_cem_best(flat, stride, a, b, bands, stats.cem_weight,
                                         stats.num_targets, score, target), pixels, threads)
    return score[:pixels], target[:pixels]


def gamma_p(a, x):
    return lib().orc_gamma_p(a, x)


def rx_threshold(bands, rate):
    v = lib().orc_rx_threshold(bands, rate)
    if v != v:
        raise ValueError("rx_threshold: bands must be positive and rate in (0, 1)")
    return v
