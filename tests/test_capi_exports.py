"""The C-ABI library loads and exports every symbol the headers declare.

CPU only: nothing here touches a GPU (the library's host-side entry points -
plane stride, limits, chi-squared threshold, synthetic frames - run without
one).
"""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from conftest import REPO

HEADERS = [os.path.join(REPO, "include", h) for h in sorted(os.listdir(os.path.join(REPO, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = set()
    for path in HEADERS:
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(hsi_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_boundary():
    names = declared_functions()
    for required in ["hsi_ctx_create", "hsi_set_library", "hsi_stats_accumulate", "hsi_stats_build",
                     "hsi_score", "hsi_score_sam_cem", "hsi_detect", "hsi_plane_stride"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    import hsi_mi355x
    lib = hsi_mi355x.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_plane_stride_pads_to_eight():
    import hsi_mi355x
    assert hsi_mi355x.plane_stride(13, 37) == 488
    assert hsi_mi355x.plane_stride(1, 1) == 8
    assert hsi_mi355x.plane_stride(128, 256) == 128 * 256
    assert hsi_mi355x.plane_stride(0, 5) == 0


def test_plane_stride_spreads_large_planes_over_channels():
    import hsi_mi355x
    # power-of-two planes get one extra 4 KiB page
    assert hsi_mi355x.plane_stride(512, 512) == 512 * 512 + 1024
    assert hsi_mi355x.plane_stride(2048, 2048) == 2048 * 2048 + 1024
    # already an odd number of pages after rounding up
    assert hsi_mi355x.plane_stride(1000, 1000) == 977 * 1024
    for h, w in [(256, 256), (1024, 1024), (2048, 1088), (1000, 1000), (333, 777), (4096, 1024)]:
        s = hsi_mi355x.plane_stride(h, w)
        assert s >= h * w and s % 1024 == 0 and (s // 1024) % 2 == 1
        assert s - h * w < 2048


def test_rx_threshold_matches_oracle_and_closed_form(oracle_lib):
    import hsi_mi355x
    for rate in (0.5, 0.1, 1e-3, 1e-6):
        # two degrees of freedom: survival exp(-t/2), so t = -2 ln(rate)
        assert abs(hsi_mi355x.rx_threshold(2, rate) - (-2 * math.log(rate))) < 1e-6 * max(1, -2 * math.log(rate))
    for bands in (2, 16, 50, 113, 224):
        for rate in (0.5, 0.1, 1e-3, 1e-6):
            assert hsi_mi355x.rx_threshold(bands, rate) == oracle_lib.rx_threshold(bands, rate)
    with pytest.raises(ValueError):
        hsi_mi355x.rx_threshold(0, 0.1)
    with pytest.raises(ValueError):
        hsi_mi355x.rx_threshold(10, 1.0)


def test_traits_mirror_the_reference():
    import hsi_mi355x as H
    assert not H.traits_of(H.Detector.SAM).needs_statistics
    assert H.traits_of(H.Detector.SAM).needs_targets
    assert H.traits_of(H.Detector.RX).needs_statistics and not H.traits_of(H.Detector.RX).needs_targets
    assert H.traits_of(H.Detector.ACE).needs_statistics and H.traits_of(H.Detector.ACE).needs_targets
    assert H.traits_of(H.Detector.CEM).needs_statistics and H.traits_of(H.Detector.CEM).needs_targets
    assert not H.traits_of(H.Detector.SAM).higher_is_better
    assert all(H.traits_of(d).higher_is_better for d in (H.Detector.RX, H.Detector.ACE, H.Detector.CEM))
    assert H.traits_of(H.Detector.RX).compute_bound and H.traits_of(H.Detector.ACE).compute_bound
    assert not H.traits_of(H.Detector.SAM).compute_bound and not H.traits_of(H.Detector.CEM).compute_bound
    assert H.parse_detector("rx") == H.Detector.RX
    with pytest.raises(ValueError):
        H.parse_detector("nope")


def test_detection_record_is_sixteen_bytes():
    import hsi_mi355x as H
    assert H.DETECTION_DTYPE.itemsize == 16
    assert C.sizeof(H.DetectionParams) == 16
    p = H.DetectionParams()
    assert abs(p.threshold - 0.10) < 1e-7 and p.nms_radius == 2 and p.max_detections == 4096


def test_synthetic_frames_are_deterministic_and_match_the_spec():
    import hsi_mi355x as H
    cfg = H.synth_config(64, 48, 113, 4)
    a, truth = H.synth_frame(cfg, 0, truth=True, threads=1)
    b = H.synth_frame(cfg, 0, threads=5)
    assert np.array_equal(a, b)                      # thread count does not matter
    assert not np.array_equal(a, H.synth_frame(cfg, 1))  # frames differ by their noise
    pixels = 64 * 48
    assert a.shape == (113, H.plane_stride(64, 48))
    assert (a[:, pixels:] == 0).all()
    assert (a[:, :pixels] >= 0).all()
    lib = H.synth_library(cfg)
    assert lib.shape == (4, 113)
    assert lib.min() >= 0.05 - 1e-6 and lib.max() <= 0.9 + 1e-6
    # each target planted in about 0.5% of pixels, labels disjoint
    counts = np.bincount(truth, minlength=256)
    assert counts[5:255].sum() == 0
    big = H.synth_config(256, 256, 16, 4)
    _, t2 = H.synth_frame(big, 0, truth=True)
    frac = np.bincount(t2, minlength=5)[1:5] / t2.size
    assert np.all(np.abs(frac - 0.005) < 0.0015)
    # a planted pixel lies closer (in angle) to its target than the background does
    x = a[:, :pixels].astype(np.float64)
    t0 = lib[0].astype(np.float64)
    cos = (t0 @ x) / (np.linalg.norm(t0) * np.linalg.norm(x, axis=0))
    if (truth == 1).any():
        assert np.median(cos[truth == 1]) > np.median(cos[truth == 0])
