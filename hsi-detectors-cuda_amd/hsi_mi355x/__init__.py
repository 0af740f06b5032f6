"""Python binding of the MI355X frame-scoring library (include/hsi_mi355x.h).

A thin ctypes layer over ``lib/libhsi_mi355x.so``: every call goes through the
C ABI. PyTorch supplies device memory and the stream (tensor ``data_ptr()``
and ``torch.cuda.current_stream().cuda_stream``); it computes nothing here.
There is no CPU fallback: if the library is not built, or no GPU is visible,
the calls raise.

Names follow the reference's host API: ``Detector`` mirrors hsi::Detector
(include/hsi/detector.hpp:23-42), ``traits_of`` mirrors DetectorTraits
(detector.hpp:44-55), ``DetectionParams`` mirrors hsi::DetectionParams
(include/hsi/types.hpp:122-140) and ``Detection`` hsi::Detection
(types.hpp:101-106).
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libhsi_mi355x.so")


class HsiError(RuntimeError):
    """A non-OK status from the C ABI, with the context's message."""

    def __init__(self, status, message):
        super().__init__(f"{status_name(status)}: {message}")
        self.status = status


OK, ERR_INVALID_ARG, ERR_CAPACITY, ERR_NOT_POSITIVE_DEFINITE, ERR_DEVICE, ERR_NO_STATISTICS, \
    ERR_NO_LIBRARY = range(7)
DETECT_NO_ALLOC_MAX_PIXELS = 1 << 24
# Statistics status-word bits (HSI_STATUS_*).
STATUS_COV_NOT_PD, STATUS_CORR_NOT_PD, STATUS_BELOW_TOLERANCE = 1, 2, 4
F32, F16, Q16 = 0, 1, 2
PLAIN, MIXED, FP32 = 0, 1, 2


class Detector(enum.IntEnum):
    SAM = 0
    RX = 1
    ACE = 2
    CEM = 3


@dataclass(frozen=True)
class DetectorTraits:
    name: str
    needs_targets: bool
    needs_statistics: bool
    higher_is_better: bool
    compute_bound: bool


_TRAITS = {
    Detector.SAM: DetectorTraits("sam", True, False, False, False),
    Detector.RX: DetectorTraits("rx", False, True, True, True),
    Detector.ACE: DetectorTraits("ace", True, True, True, True),
    Detector.CEM: DetectorTraits("cem", True, True, True, False),
}


def traits_of(detector):
    return _TRAITS[Detector(detector)]


def parse_detector(name):
    for d, t in _TRAITS.items():
        if t.name == name:
            return d
    raise ValueError(f"unknown detector '{name}' (sam, rx, ace, cem)")


class _Detection(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("target", C.c_int32), ("score", C.c_float)]


DETECTION_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("target", "<i4"), ("score", "<f4")])


class DetectionParams(C.Structure):
    """hsi_detection_params; defaults as hsi::DetectionParams."""

    _fields_ = [("threshold", C.c_float), ("nms_radius", C.c_int32),
                ("max_detections", C.c_int32), ("higher_is_better", C.c_int32)]

    def __init__(self, threshold=0.10, nms_radius=2, max_detections=4096, higher_is_better=False):
        super().__init__(threshold, nms_radius, max_detections, int(bool(higher_is_better)))


class MlpTrainParams(C.Structure):
    """hsi_mlp_train_params (Adam + positive-class weight)."""

    _fields_ = [("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("epsilon", C.c_float), ("pos_weight", C.c_float)]

    def __init__(self, learning_rate=1e-3, beta1=0.9, beta2=0.999, epsilon=1e-8, pos_weight=1.0):
        super().__init__(learning_rate, beta1, beta2, epsilon, pos_weight)


class EvalResult(C.Structure):
    """hsi_eval_result."""
    _fields_ = [("auc", C.c_double), ("positives", C.c_uint64), ("negatives", C.c_uint64),
                ("tp", C.c_uint64), ("fp", C.c_uint64), ("fn", C.c_uint64), ("tn", C.c_uint64)]


class Scene(C.Structure):
    """hsi_scene: one member of a batch (device cube, host library, output maps)."""

    _fields_ = [("d_cube", C.c_void_p), ("height", C.c_int), ("width", C.c_int), ("bands", C.c_int),
                ("plane_stride", C.c_size_t), ("library", C.c_void_p), ("num_targets", C.c_int),
                ("d_sam_score", C.c_void_p), ("d_sam_target", C.c_void_p),
                ("d_cem_score", C.c_void_p), ("d_cem_target", C.c_void_p),
                ("d_rx_score", C.c_void_p), ("d_ace_score", C.c_void_p), ("d_ace_target", C.c_void_p)]


class SynthConfig(C.Structure):
    _fields_ = [("height", C.c_int32), ("width", C.c_int32), ("bands", C.c_int32),
                ("num_targets", C.c_int32), ("num_endmembers", C.c_int32),
                ("drift_period", C.c_int32), ("plant_fraction", C.c_double),
                ("anomaly_fraction", C.c_double), ("noise_sigma", C.c_double),
                ("drift", C.c_double), ("seed", C.c_uint64)]


_lib = None
_vp = C.c_void_p


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                           f"(or `make -C hsi-detectors-cuda_amd`)")
    # torch carries its own HIP runtime (same soname as ROCm's, a different
    # file). Loading torch first lets this library bind to that one copy;
    # loading it first would put a second HIP runtime in the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    sig = {
        "hsi_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
        "hsi_ctx_destroy": (None, [_vp]),
        "hsi_ctx_last_error": (C.c_char_p, [_vp]),
        "hsi_status_name": (C.c_char_p, [C.c_int]),
        "hsi_ctx_reserve": (C.c_int, [_vp, C.c_size_t, C.c_int]),
        "hsi_plane_stride": (C.c_size_t, [C.c_int, C.c_int]),
        "hsi_max_targets": (C.c_int, []),
        "hsi_max_bands": (C.c_int, []),
        "hsi_set_library": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int]),
        "hsi_stats_accumulate_screen": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_size_t, _vp,
                                                  C.c_float, _vp, _vp, C.c_int, _vp, C.c_size_t, _vp]),
        "hsi_stats_accumulate": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_size_t, C.c_int, _vp, _vp]),
        "hsi_encode_q16": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_size_t, _vp, _vp]),
        "hsi_set_q16_code": (C.c_int, [_vp, C.c_int, _vp, _vp]),
        "hsi_get_q16_code": (C.c_int, [_vp, C.c_int, _vp, _vp]),
        "hsi_stats_build": (C.c_int, [_vp, C.c_double, _vp]),
        "hsi_stats_check": (C.c_int, [_vp, _vp]),
        "hsi_stats_read": (C.c_int, [_vp] + [_vp] * 8),
        "hsi_stats_upload": (C.c_int, [_vp, C.c_int] + [_vp] * 5),
        "hsi_score": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_size_t, _vp, _vp, _vp, _vp]),
        "hsi_score_sam_cem": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_size_t,
                                        _vp, _vp, _vp, _vp, _vp]),
        "hsi_detect": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, DetectionParams, _vp, _vp, _vp]),
        "hsi_detect_no_alloc": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, DetectionParams, _vp, _vp,
                                          _vp]),
        "hsi_narrow_to_half": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
        "hsi_transpose_bip_to_bsq": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_size_t,
                                               _vp]),
        "hsi_rx_threshold": (C.c_double, [C.c_int, C.c_double]),
        "hsi_stats_status_copy": (C.c_int, [_vp, _vp, _vp]),
        "hsi_mlp_param_count": (C.c_size_t, [C.c_int]),
        "hsi_mlp_init": (C.c_int, [_vp, C.c_int, C.c_uint64]),
        "hsi_mlp_set_weights": (C.c_int, [_vp, C.c_int, _vp]),
        "hsi_mlp_get_weights": (C.c_int, [_vp, _vp]),
        "hsi_mlp_gradients": (C.c_int, [_vp, _vp, C.c_size_t, C.c_int, _vp, C.c_float, _vp, _vp,
                                        _vp]),
        "hsi_mlp_train_step": (C.c_int, [_vp, _vp, C.c_size_t, C.c_int, _vp,
                                         C.POINTER(MlpTrainParams), _vp, _vp]),
        "hsi_mlp_infer": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_size_t,
                                    C.c_int, _vp, _vp]),
        "hsi_gather_pixels": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_size_t, _vp,
                                        C.c_int, C.c_int, _vp, C.c_size_t, _vp]),
        "hsi_score_candidates": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int,
                                           C.c_size_t, _vp, _vp, C.c_int, _vp, _vp, _vp, _vp]),
        "hsi_mlp_screen": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_float,
                                     _vp, _vp, C.c_int, _vp, C.c_size_t, _vp, _vp]),
        "hsi_score_candidate_spectra": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int,
                                                  C.c_size_t, _vp, _vp, C.c_int, _vp, C.c_size_t,
                                                  _vp, _vp, _vp, _vp]),
        "hsi_evaluate": (C.c_int, [_vp, _vp, _vp, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_float,
                                   C.POINTER(EvalResult), _vp, _vp]),
        "hsi_demosaic": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp,
                                   C.c_size_t, _vp]),
        "hsi_batch_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
        "hsi_batch_destroy": (None, [_vp]),
        "hsi_batch_last_error": (C.c_char_p, [_vp]),
        "hsi_batch_set_scenes": (C.c_int, [_vp, C.POINTER(Scene), C.c_int]),
        "hsi_batch_statistics": (C.c_int, [_vp, C.c_double, _vp]),
        "hsi_batch_check": (C.c_int, [_vp, C.POINTER(C.c_int)]),
        "hsi_batch_score": (C.c_int, [_vp, _vp]),
        "hsi_batch_run": (C.c_int, [_vp, C.c_double, C.c_int, _vp]),
        "hsi_batch_autotune": (C.c_int, [_vp, _vp, C.c_char_p, C.c_size_t]),
        "hsi_batch_context": (_vp, [_vp, C.c_int]),
        "hsi_synth_default": (None, [C.POINTER(SynthConfig)]),
        "hsi_synth_library": (C.c_int, [C.POINTER(SynthConfig), _vp]),
        "hsi_synth_anomaly": (C.c_int, [C.POINTER(SynthConfig), _vp]),
        "hsi_synth_frame": (C.c_int, [C.POINTER(SynthConfig), C.c_uint64, _vp, C.c_size_t, _vp,
                                      C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def status_name(status):
    return lib().hsi_status_name(status).decode()


def plane_stride(height, width):
    return int(lib().hsi_plane_stride(height, width))


def rx_threshold(bands, rate):
    v = lib().hsi_rx_threshold(bands, rate)
    if v != v:
        raise ValueError("rx_threshold: bands must be positive and rate in (0, 1)")
    return v


# ---- synthetic frames ---------------------------------------------------------

def synth_config(height=256, width=256, bands=113, num_targets=4, **kw):
    cfg = SynthConfig()
    lib().hsi_synth_default(C.byref(cfg))
    cfg.height, cfg.width, cfg.bands, cfg.num_targets = height, width, bands, num_targets
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def synth_library(cfg):
    out = np.empty((cfg.num_targets, cfg.bands), np.float32)
    if lib().hsi_synth_library(C.byref(cfg), out.ctypes.data) != 0:
        raise ValueError("bad synthetic configuration")
    return out


def synth_anomaly(cfg):
    out = np.empty(cfg.bands, np.float32)
    if lib().hsi_synth_anomaly(C.byref(cfg), out.ctypes.data) != 0:
        raise ValueError("bad synthetic configuration")
    return out


def synth_frame(cfg, frame=0, stride=None, truth=False, threads=0, out=None):
    """BSQ float32 frame as a (bands, stride) array; optionally the truth mask."""
    pixels = cfg.height * cfg.width
    stride = plane_stride(cfg.height, cfg.width) if stride is None else stride
    cube = np.empty((cfg.bands, stride), np.float32) if out is None else out
    mask = np.empty(pixels, np.uint8) if truth else None
    rc = lib().hsi_synth_frame(C.byref(cfg), frame, cube.ctypes.data, stride,
                               None if mask is None else mask.ctypes.data, threads)
    if rc != 0:
        raise ValueError("bad synthetic configuration")
    return (cube, mask) if truth else cube


# ---- device helpers -------------------------------------------------------------

def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the MI355X library has no CPU path")
    return torch


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        return C.c_void_p(_torch().cuda.current_stream().cuda_stream)
    return C.c_void_p(getattr(stream, "cuda_stream", stream))


def _dtype_of(cube):
    """hsi_dtype of a device tensor: float32 -> F32, float16 -> F16, a 16-bit
    integer tensor -> Q16 (the codes of hsi_encode_q16)."""
    d = str(cube.dtype)
    if d == "torch.float16":
        return F16
    if d in ("torch.int16", "torch.uint16"):
        return Q16
    return F32


def q16_decode(codes, offset, scale):
    """The device's Q16 decode, fmaf(scale[b], q, offset[b]) rounded once to
    float32, on the host: codes (bands, n) uint16 -> float32. The product of a
    float and a 16-bit integer and its sum with a float are exact in float64
    here, so one rounding to float32 matches the fused multiply-add."""
    q = np.asarray(codes).view(np.uint16).astype(np.float64)
    off = np.asarray(offset, np.float32).astype(np.float64)[:, None]
    sc = np.asarray(scale, np.float32).astype(np.float64)[:, None]
    return (sc * q + off).astype(np.float32)


def to_device_cube(cube, height, width, dtype="f32", device="cuda"):
    """Host BSQ cube (bands, H, W) or (bands, >=H*W) -> device (bands, stride) tensor."""
    torch = _torch()
    cube = np.asarray(cube)
    bands = cube.shape[0]
    flat = cube.reshape(bands, -1)
    pixels = height * width
    stride = plane_stride(height, width)
    host = np.zeros((bands, stride), np.float32)
    host[:, :pixels] = flat[:, :pixels]
    t = torch.from_numpy(host).to(device)
    if dtype == "f16":
        t = t.half()
    return t


class Context:
    """hsi_ctx: a device, a target library and scene statistics."""

    def __init__(self, device=0):
        _torch()
        self._lib = lib()
        h = C.c_void_p()
        st = self._lib.hsi_ctx_create(device, C.byref(h))
        if st != OK:
            raise HsiError(st, f"cannot create a context on device {device}")
        self.handle = h
        self.device = device

    def close(self):
        if self.handle and not getattr(self, "_borrowed", False):
            self._lib.hsi_ctx_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != OK:
            raise HsiError(st, self._lib.hsi_ctx_last_error(self.handle).decode())

    def reserve(self, max_pixels, max_bands):
        self._check(self._lib.hsi_ctx_reserve(self.handle, max_pixels, max_bands))

    def set_library(self, values, norms=None):
        values = np.ascontiguousarray(values, np.float32)
        t, b = values.shape
        nptr = None
        if norms is not None:
            norms = np.ascontiguousarray(norms, np.float32)
            nptr = norms.ctypes.data
        self._check(self._lib.hsi_set_library(self.handle, values.ctypes.data, nptr, t, b))
        self.num_targets, self.bands = t, b

    # statistics
    def stats_accumulate(self, cube, height, width, bands, precision=PLAIN, shift=None, stream=None):
        """Moments of a device cube (float32, float16, or a Q16 code as
        int16/uint16); `shift` (device float32 [bands]) is the MIXED form's
        reference, the frame's sampled mean when None."""
        self._check(self._lib.hsi_stats_accumulate(self.handle, _ptr(cube), _dtype_of(cube), height, width,
                                                   bands, cube.stride(0), precision, _ptr(shift),
                                                   _stream(stream)))

    def stats_accumulate_screen(self, cube, height, width, bands, cut, candidates, count, spectra=None,
                                shift=None, stream=None):
        """MIXED moments of a float32 cube and, from the same read, the
        pre-screen's candidates at `cut` (records in `candidates` [cap][4]
        int32, rows in `spectra` [cap][pitch] float32 when given); the
        network sees the pixel minus `shift` (the sampled mean when None)."""
        cap = candidates.shape[0]
        if spectra is not None and spectra.shape[0] < cap:
            raise ValueError(f"spectra hold {spectra.shape[0]} rows; the list's capacity is {cap}")
        self._check(self._lib.hsi_stats_accumulate_screen(
            self.handle, _ptr(cube), height, width, bands, cube.stride(0), _ptr(shift), float(cut),
            _ptr(candidates), _ptr(count), cap, _ptr(spectra), spectra.stride(0) if spectra is not None else 0,
            _stream(stream)))

    def encode_q16(self, src, height, width, bands, dst, stream=None):
        """hsi_encode_q16: device float32 (bands, stride) cube -> Q16 codes in
        `dst` (int16 or uint16 tensor of the same shape); the context keeps
        the code."""
        self._check(self._lib.hsi_encode_q16(self.handle, _ptr(src), height, width, bands, src.stride(0),
                                             _ptr(dst), _stream(stream)))

    def set_q16_code(self, offset, scale):
        offset = np.ascontiguousarray(offset, np.float32)
        scale = np.ascontiguousarray(scale, np.float32)
        self._check(self._lib.hsi_set_q16_code(self.handle, offset.size, offset.ctypes.data,
                                               scale.ctypes.data))

    def q16_code(self, bands):
        """(offset, scale) float32 arrays of the context's Q16 code."""
        off = np.zeros(bands, np.float32)
        sc = np.zeros(bands, np.float32)
        self._check(self._lib.hsi_get_q16_code(self.handle, bands, off.ctypes.data, sc.ctypes.data))
        return off, sc

    def stats_build(self, loading, stream=None):
        self._check(self._lib.hsi_stats_build(self.handle, loading, _stream(stream)))

    def stats_check(self, stream=None):
        self._check(self._lib.hsi_stats_check(self.handle, _stream(stream)))

    def stats_status(self, stream=None):
        """The two status words of the last build (covariance, autocorrelation
        path) as ints: 0, or STATUS_* bits."""
        torch = _torch()
        words = torch.zeros(2, dtype=torch.int32, device="cuda")
        self._check(self._lib.hsi_stats_status_copy(self.handle, _ptr(words), _stream(stream)))
        torch.cuda.synchronize()
        return [int(v) for v in words.cpu().tolist()]

    def stats_read(self, bands, num_targets):
        out = {
            "mean": np.zeros(bands, np.float32),
            "whitener": np.zeros((bands, bands), np.float32),
            "target_white": np.zeros((max(num_targets, 1), bands), np.float32),
            "target_norm2": np.zeros(max(num_targets, 1), np.float32),
            "cem_weight": np.zeros((max(num_targets, 1), bands), np.float32),
            "count": np.zeros(1, np.float64),
            "sum": np.zeros(bands, np.float64),
            "sum_outer": np.zeros((bands, bands), np.float64),
        }
        self._check(self._lib.hsi_stats_read(self.handle, *[out[k].ctypes.data for k in (
            "mean", "whitener", "target_white", "target_norm2", "cem_weight", "count", "sum",
            "sum_outer")]))
        return out

    def moments_read(self, bands):
        """Raw FP64 moments of the last accumulation: count, sum[B] and
        sum(x x^T) as [B][B] (lower triangle valid)."""
        out = {"count": np.zeros(1, np.float64), "sum": np.zeros(bands, np.float64),
               "sum_outer": np.zeros((bands, bands), np.float64)}
        self._check(self._lib.hsi_stats_read(self.handle, None, None, None, None, None,
                                             out["count"].ctypes.data, out["sum"].ctypes.data,
                                             out["sum_outer"].ctypes.data))
        return out

    def stats_upload(self, bands, mean, whitener, target_white=None, target_norm2=None,
                     cem_weight=None):
        def p(a):
            return None if a is None else np.ascontiguousarray(a, np.float32)
        arrs = [p(mean), p(whitener), p(target_white), p(target_norm2), p(cem_weight)]
        self._keep = arrs
        self._check(self._lib.hsi_stats_upload(
            self.handle, bands, *[None if a is None else a.ctypes.data for a in arrs]))

    # scoring
    def score(self, detector, cube, height, width, bands, score, target=None, aux=None,
              precision=PLAIN, stream=None):
        dtype = _dtype_of(cube)
        self._check(self._lib.hsi_score(self.handle, int(detector), precision, _ptr(cube), dtype,
                                        height, width, bands, cube.stride(0), _ptr(score),
                                        _ptr(target), _ptr(aux), _stream(stream)))

    def score_sam_cem(self, cube, height, width, bands, sam_score=None, sam_target=None,
                      cem_score=None, cem_target=None, stream=None):
        dtype = _dtype_of(cube)
        self._check(self._lib.hsi_score_sam_cem(self.handle, _ptr(cube), dtype, height, width,
                                                bands, cube.stride(0), _ptr(sam_score),
                                                _ptr(sam_target), _ptr(cem_score),
                                                _ptr(cem_target), _stream(stream)))

    def detect(self, score, target, height, width, params, out, count, stream=None):
        self._check(self._lib.hsi_detect(self.handle, _ptr(score), _ptr(target), height, width,
                                         params, _ptr(out), _ptr(count), _stream(stream)))

    def detect_no_alloc(self, score, target, height, width, params, out, count, stream=None):
        """hsi_detect_no_alloc: hsi_detect over a static per-GPU scratch
        (frames up to DETECT_NO_ALLOC_MAX_PIXELS; overlapping calls race)."""
        self._check(self._lib.hsi_detect_no_alloc(self.handle, _ptr(score), _ptr(target), height,
                                                  width, params, _ptr(out), _ptr(count),
                                                  _stream(stream)))

    # learned pre-screen
    def mlp_init(self, bands, seed=1):
        self._check(self._lib.hsi_mlp_init(self.handle, bands, seed))
        self.mlp_bands = bands

    def mlp_set_weights(self, bands, weights):
        w = np.ascontiguousarray(weights, np.float32)
        assert w.size == mlp_param_count(bands)
        self._check(self._lib.hsi_mlp_set_weights(self.handle, bands, w.ctypes.data))
        self.mlp_bands = bands

    def mlp_get_weights(self):
        w = np.empty(mlp_param_count(self.mlp_bands), np.float32)
        self._check(self._lib.hsi_mlp_get_weights(self.handle, w.ctypes.data))
        return w

    def mlp_gradients(self, x, labels, pos_weight, grad, loss=None, n=None, stream=None):
        """x: device (bands, >= n) float32 tensor, band-major samples."""
        n = labels.numel() if n is None else n
        self._check(self._lib.hsi_mlp_gradients(self.handle, _ptr(x), x.stride(0), n, _ptr(labels),
                                                pos_weight, _ptr(grad), _ptr(loss),
                                                _stream(stream)))

    def mlp_train_step(self, x, labels, params, loss=None, n=None, stream=None):
        n = labels.numel() if n is None else n
        self._check(self._lib.hsi_mlp_train_step(self.handle, _ptr(x), x.stride(0), n, _ptr(labels),
                                                 C.byref(params), _ptr(loss), _stream(stream)))

    def mlp_infer(self, cube, height, width, bands, logit, center=True, stream=None):
        dtype = _dtype_of(cube)
        self._check(self._lib.hsi_mlp_infer(self.handle, _ptr(cube), dtype, height, width, bands,
                                            cube.stride(0), int(bool(center)), _ptr(logit),
                                            _stream(stream)))

    def gather_pixels(self, cube, height, width, bands, index, out, center=True, stream=None):
        n = index.numel()
        self._check(self._lib.hsi_gather_pixels(self.handle, _ptr(cube), height, width, bands,
                                                cube.stride(0), _ptr(index), n, int(bool(center)),
                                                _ptr(out), out.stride(0), _stream(stream)))

    def score_candidates(self, detector, cube, height, width, bands, candidates, count, score,
                         target=None, aux=None, stream=None):
        cap = candidates.shape[0]
        self._check(self._lib.hsi_score_candidates(self.handle, int(detector), _ptr(cube), height,
                                                   width, bands, cube.stride(0), _ptr(candidates),
                                                   _ptr(count), cap, _ptr(score), _ptr(target),
                                                   _ptr(aux), _stream(stream)))


    def mlp_screen(self, cube, height, width, bands, cut, candidates, count, spectra=None,
                   logit=None, stream=None):
        """One pass: candidates at the cut (fixed order, not raster) and with
        `spectra` ([capacity, >= 32 * ceil(bands / 32)] float32) their bands,
        each in the row its record's target field names."""
        cap = candidates.shape[0]
        if spectra is not None and spectra.shape[0] < cap:
            raise ValueError(f"spectra hold {spectra.shape[0]} rows; the list's capacity is {cap}")
        self._check(self._lib.hsi_mlp_screen(self.handle, _ptr(cube), height, width, bands,
                                             cube.stride(0), float(cut), _ptr(candidates), _ptr(count),
                                             cap, _ptr(spectra), spectra.stride(0) if spectra is not None else 0,
                                             _ptr(logit), _stream(stream)))

    def score_candidate_spectra(self, detector, cube, height, width, bands, candidates, count, spectra,
                                score, target=None, aux=None, precision=PLAIN, stream=None):
        cap = candidates.shape[0]
        if spectra.shape[0] < cap:
            raise ValueError(f"spectra hold {spectra.shape[0]} rows; the list's capacity is {cap}")
        self._check(self._lib.hsi_score_candidate_spectra(
            self.handle, int(detector), int(precision), _ptr(cube), height, width, bands, cube.stride(0),
            _ptr(candidates), _ptr(count), cap, _ptr(spectra), spectra.stride(0), _ptr(score),
            _ptr(target), _ptr(aux), _stream(stream)))


    def evaluate(self, score, labels, target_class=1, max_label=8, higher_is_better=False,
                 threshold=0.10, stream=None):
        """hsi_evaluate on device tensors (float32 scores, uint8 labels):
        AUC, positives/negatives, tp/fp/fn/tn and the per-class table
        {label: (count, mean, min)} for labels present."""
        n = score.numel()
        if labels.numel() != n:
            raise ValueError("score map and labels differ in size")
        res = EvalResult()
        stats = np.zeros(3 * (max_label + 1), np.float64)
        self._check(self._lib.hsi_evaluate(self.handle, _ptr(score), _ptr(labels), n, target_class,
                                           max_label, int(bool(higher_is_better)), float(threshold),
                                           C.byref(res), stats.ctypes.data, _stream(stream)))
        out = {k: getattr(res, k) for k, _ in EvalResult._fields_}
        out["classes"] = {c: (int(stats[3 * c]), float(stats[3 * c + 1]), float(stats[3 * c + 2]))
                          for c in range(max_label + 1) if stats[3 * c] > 0}
        return out


RAW_U8, RAW_U16, RAW_F32 = 0, 1, 2


class Batch:
    """hsi_batch: uneven scenes scored by one persistent launch per shape class."""

    def __init__(self, device=0):
        _torch()
        self._lib = lib()
        h = C.c_void_p()
        st = self._lib.hsi_batch_create(device, C.byref(h))
        if st != OK:
            raise HsiError(st, f"cannot create a batch on device {device}")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self._lib.hsi_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != OK:
            raise HsiError(st, self._lib.hsi_batch_last_error(self.handle).decode())

    def set_scenes(self, scenes):
        """scenes: list of dicts with cube (device (bands, stride) f32 tensor),
        height, width, library (host [T, B] or None) and optional output
        tensors sam_score, sam_target, cem_score, cem_target, rx, ace_score,
        ace_target."""
        arr = (Scene * len(scenes))()
        self._keep = []
        for i, d in enumerate(scenes):
            cube = d["cube"]
            libv = d.get("library")
            if libv is not None:
                libv = np.ascontiguousarray(libv, np.float32)
                self._keep.append(libv)
            arr[i] = Scene(cube.data_ptr(), d["height"], d["width"], cube.shape[0], cube.stride(0),
                           None if libv is None else libv.ctypes.data, 0 if libv is None else libv.shape[0],
                           *[None if d.get(k) is None else d[k].data_ptr() for k in
                             ("sam_score", "sam_target", "cem_score", "cem_target", "rx", "ace_score",
                              "ace_target")])
        self._check(self._lib.hsi_batch_set_scenes(self.handle, arr, len(scenes)))

    def statistics(self, loading=1e-6, stream=None):
        self._check(self._lib.hsi_batch_statistics(self.handle, loading, _stream(stream)))

    def check(self):
        idx = C.c_int(-1)
        self._check(self._lib.hsi_batch_check(self.handle, C.byref(idx)))

    def score(self, stream=None):
        self._check(self._lib.hsi_batch_score(self.handle, _stream(stream)))

    def run(self, loading=1e-6, use_graph=True, stream=None):
        """Statistics and scoring of every scene; with use_graph a captured
        hipGraph replayed per call."""
        self._check(self._lib.hsi_batch_run(self.handle, loading, int(bool(use_graph)), _stream(stream)))

    def context(self, scene):
        """The scene's context (owned by the batch) as a Context."""
        h = self._lib.hsi_batch_context(self.handle, scene)
        if not h:
            raise IndexError(scene)
        c = Context.__new__(Context)
        c._lib = self._lib
        c.handle = C.c_void_p(h)
        c._borrowed = True
        return c

    def autotune(self, stream=None):
        buf = C.create_string_buffer(8192)
        self._check(self._lib.hsi_batch_autotune(self.handle, _stream(stream), buf, len(buf)))
        return buf.value.decode()


def demosaic(ctx, raw, raw_type, raw_height, raw_width, period, bands, cube, band_order=None, stream=None):
    """Snapshot-mosaic frame (device tensor) -> BSQ cube tensor (bands, stride)."""
    order = None
    if band_order is not None:
        order = np.ascontiguousarray(band_order, np.int32)
    ctx._check(ctx._lib.hsi_demosaic(ctx.handle, _ptr(raw), raw_type, raw_height, raw_width, period, bands,
                                     None if order is None else order.ctypes.data, _ptr(cube),
                                     cube.stride(0), _stream(stream)))


def narrow_to_half(src, dst, count=None, stream=None):
    """hsi_narrow_to_half: `count` float32 elements of device tensor `src`
    to float16 in `dst` (both 16-byte aligned)."""
    count = src.numel() if count is None else count
    st = lib().hsi_narrow_to_half(_ptr(src), _ptr(dst), count, _stream(stream))
    if st != OK:
        raise HsiError(st, "narrow_to_half: null or misaligned buffer")


def transpose_bip_to_bsq(src, dst, height, width, bands, stream=None):
    """hsi_transpose_bip_to_bsq: device BIP [pixel][band] float32 -> BSQ
    (bands, stride) tensor `dst`."""
    st = lib().hsi_transpose_bip_to_bsq(_ptr(src), _ptr(dst), height, width, bands,
                                        dst.stride(0), _stream(stream))
    if st != OK:
        raise HsiError(st, "transpose_bip_to_bsq: bad shape or stride")


def mlp_param_count(bands):
    return int(lib().hsi_mlp_param_count(bands))


def detections_to_numpy(out_tensor, count):
    """Device detection buffer (int32 [cap, 4]) -> structured array of `count` rows."""
    raw = out_tensor[:count].cpu().numpy()
    return raw.view(np.uint8).reshape(-1, 16).copy().view(DETECTION_DTYPE).reshape(-1)
