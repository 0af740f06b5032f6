"""The precision form's storage: a 16-bit per-band affine code (Q16) of an
FP32 frame, written on the device, read by the mixed statistics and the
mixed RX/ACE kernels.

Bars (the precision form's, stated per detector): RX relative error <= 1e-3
(floor 1), ACE absolute error <= 1e-3, against FP64 scores; detection recall
>= 0.999 against the plain lists.
  (a) kernel error: the oracle scores the values the code decodes to, with
      statistics from those values, so the error is the arithmetic's alone;
  (b) end to end: the oracle scores the FP32 frame itself, with statistics
      from it - the error of storing the frame in 16 bits included.
"""
import os

import numpy as np
import pytest

from gpu_util import rel_err

pytestmark = pytest.mark.gpu

BAR = 1e-3
LOADING = 1e-6
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def H():
    import hsi_mi355x
    return hsi_mi355x


def encode(H, torch, ctx, host, h, w, b):
    src = torch.from_numpy(np.ascontiguousarray(host)).cuda()
    codes = torch.empty(src.shape, dtype=torch.int16, device="cuda")
    ctx.encode_q16(src, h, w, b, codes)
    torch.cuda.synchronize()
    return src, codes


def test_q16_encoding_bounds_and_code(H, torch):
    """Codes decode to within half a code step of the input; the code is the
    band's minimum and range / 65535; a constant band codes to 0 with scale
    0; plane padding codes to 0."""
    h, w, b = 37, 51, 19
    rng = np.random.default_rng(5)
    stride = H.plane_stride(h, w)
    host = np.zeros((b, stride), np.float32)
    host[:, :h * w] = rng.normal(0.4, 0.2, (b, h * w)) * np.exp(rng.uniform(-3, 3, (b, 1)))
    host[3, :h * w] = 0.625            # a constant band
    host[:, h * w:] = 123.0            # padding that must not count
    ctx = H.Context(0)
    _, codes = encode(H, torch, ctx, host, h, w, b)
    q = codes.cpu().numpy().view(np.uint16)
    off, sc = ctx.q16_code(b)
    x = host[:, :h * w]
    assert np.array_equal(off, x.min(axis=1))
    want_sc = ((x.max(axis=1) - x.min(axis=1)) / np.float32(65535)).astype(np.float32)
    assert np.array_equal(sc, want_sc)
    assert sc[3] == 0 and not q[3].any()
    assert not q[:, h * w:].any()
    dec = H.q16_decode(q[:, :h * w], off, sc).astype(np.float64)
    step = sc.astype(np.float64)[:, None]
    # a code within half a step, plus the FP32 rounding of (x - offset) * (65535 / range)
    assert np.all(np.abs(dec - x) <= 0.53 * step + 2e-7 * np.abs(x))
    assert q[:, :h * w].max() == 65535 and q[np.arange(b) != 3][:, :h * w].min() == 0
    ctx.close()


def test_q16_requires_a_code_and_mixed_forms(H, torch):
    ctx = H.Context(0)
    codes = torch.zeros((8, 64), dtype=torch.int16, device="cuda")
    with pytest.raises(H.HsiError, match="Q16 code"):
        ctx.stats_accumulate(codes, 8, 8, 8, precision=H.MIXED)
    ctx.set_q16_code(np.zeros(8, np.float32), np.ones(8, np.float32))
    with pytest.raises(H.HsiError, match="F32"):
        ctx.stats_accumulate(codes, 8, 8, 8, precision=H.PLAIN)
    ctx.close()


@pytest.mark.parametrize("shift", ["sampled", "given"])
def test_mixed_moments_of_16_bit_cubes(H, torch, oracle_lib, shift):
    """Mixed moments of an F16 and a Q16 cube against the oracle's exact
    moments of the values those cubes hold, with the sampled reference and
    with a caller's reference (another frame's mean, the lagged mode)."""
    h, w, b, t = 64, 96, 113, 4
    cfg = H.synth_config(h, w, b, t, seed=31)
    host = H.synth_frame(cfg, 0)
    p = h * w
    ctx = H.Context(0)
    src, codes = encode(H, torch, ctx, host, h, w, b)
    off, sc = ctx.q16_code(b)
    given = None
    if shift == "given":
        other = H.synth_frame(cfg, 7)[:, :p].astype(np.float64).mean(axis=1).astype(np.float32)
        given = torch.from_numpy(other).cuda()
    for name, cube, values in (("f16", src.half(), host[:, :p].astype(np.float16).astype(np.float32)),
                               ("q16", codes, H.q16_decode(codes.cpu().numpy()[:, :p], off, sc))):
        ctx.stats_accumulate(cube, h, w, b, precision=H.MIXED, shift=given)
        got = ctx.moments_read(b)
        cnt, s1, s2 = oracle_lib.moments(np.ascontiguousarray(values).reshape(b, h, w))
        assert got["count"][0] == cnt
        e1 = np.abs(got["sum"] - s1).max() / np.abs(s1).max()
        lo = np.tril_indices(b)
        e2 = np.abs(got["sum_outer"][lo] - s2[lo]).max() / np.abs(s2[lo]).max()
        print(f"{name} {shift}: sum rel {e1:.2e}, outer rel {e2:.2e}")
        assert e1 <= 1e-6 and e2 <= 1e-6
    ctx.close()


def scores_q16(H, torch, host, lib, h, w, b, tiles=None):
    """The precision path on the device: encode, mixed moments of the codes,
    FP64 build, mixed ACE with the RX side output."""
    p = h * w
    ctx = H.Context(0)
    ctx.set_library(lib)
    src, codes = encode(H, torch, ctx, host, h, w, b)
    del src
    ctx.stats_accumulate(codes, h, w, b, precision=H.MIXED)
    ctx.stats_build(LOADING)
    ctx.stats_check()
    ace = torch.empty(p, dtype=torch.float32, device="cuda")
    rx = torch.empty(p, dtype=torch.float32, device="cuda")
    tg = torch.empty(p, dtype=torch.int32, device="cuda")
    ctx.score(H.Detector.ACE, codes, h, w, b, ace, tg, rx, precision=H.MIXED)
    rx_only = torch.empty(p, dtype=torch.float32, device="cuda")
    ctx.score(H.Detector.RX, codes, h, w, b, rx_only, precision=H.MIXED)
    torch.cuda.synchronize()
    off, sc = ctx.q16_code(b)
    q = codes.cpu().numpy()[:, :p]
    out = dict(ace=ace.cpu().numpy(), rx=rx.cpu().numpy(), target=tg.cpu().numpy(),
               rx_only=rx_only.cpu().numpy(), decoded=H.q16_decode(q, off, sc))
    ctx.close()
    return out


@pytest.mark.parametrize("name", ["scene_32x40x113_t4", "scene_20x20x224_t16"])
def test_q16_kernel_error_on_fixture_scenes(H, torch, oracle_lib, name):
    """(a) on the golden scenes: the oracle fed the decoded values."""
    from conftest import golden
    g = golden(name)
    b, h, w = g["cube"].shape
    host = np.zeros((b, H.plane_stride(h, w)), np.float32)
    host[:, :h * w] = g["cube"].reshape(b, -1)
    got = scores_q16(H, torch, host, g["library"], h, w, b)
    dec = np.ascontiguousarray(got["decoded"]).reshape(b, h, w)
    st = oracle_lib.statistics_of_cube(dec, g["library"], LOADING)
    assert st.valid, st.error
    want_ace, want_t, want_rx = oracle_lib.ace_best(dec, st)
    e_rx = rel_err(got["rx"], want_rx, 1.0).max()
    e_ace = np.abs(got["ace"].astype(np.float64) - want_ace).max()
    print(f"{name}: (a) RX max rel {e_rx:.2e}, ACE max abs {e_ace:.2e}")
    assert e_rx <= BAR and e_ace <= BAR
    assert np.array_equal(got["rx"], got["rx_only"])


def recall(got_score, want_score, h, w, thr, oracle_lib, target_got=None, target_want=None):
    gl, _ = oracle_lib.detect(got_score, target_got, h, w, thr, 2, True)
    wl, wn = oracle_lib.detect(want_score, target_want, h, w, thr, 2, True)
    gset = set((gl["y"] * w + gl["x"]).tolist())
    wset = set((wl["y"] * w + wl["x"]).tolist())
    return len(gset & wset) / max(len(wset), 1), len(wset)


def test_q16_stripe_end_to_end_against_fp32(H, torch, oracle_lib):
    """(b) on a 256 x 1024 x 224 stripe with 16 targets (a quarter of the
    1024 x 1024 x 224 precision workload, at its real row length): the
    device stores only the 16-bit code; the oracle works from the FP32
    frame."""
    h, w, b, t = 256, 1024, 224, 16
    cfg = H.synth_config(h, w, b, t, seed=20261203 + 3000)
    lib = H.synth_library(cfg)
    host = H.synth_frame(cfg, 0, threads=THREADS)
    p = h * w
    print(f"Q16 stripe {h}x{w}x{b}: device path", flush=True)
    got = scores_q16(H, torch, host, lib, h, w, b)
    print("oracle on the FP32 frame", flush=True)
    cube = np.ascontiguousarray(host[:, :p]).reshape(b, h, w)
    cnt, s1, s2 = oracle_lib.moments(cube, pixels=p, threads=THREADS)
    st = oracle_lib.build_statistics(cnt, s1, s2, lib, LOADING)
    assert st.valid, st.error
    want_ace, want_t, want_rx = oracle_lib.ace_best(cube, st, pixels=p, threads=THREADS)
    e_rx = rel_err(got["rx"], want_rx, 1.0).max()
    e_ace = np.abs(got["ace"].astype(np.float64) - want_ace).max()
    agree = (got["target"] == want_t).mean()
    rx_thr = float(np.float32(oracle_lib.rx_threshold(b, 1e-3)))
    r_rx, n_rx = recall(got["rx"], want_rx, h, w, rx_thr, oracle_lib)
    r_ace, n_ace = recall(got["ace"], want_ace, h, w, 0.5, oracle_lib, got["target"], want_t)
    print(f"(b) RX max rel {e_rx:.2e}, ACE max abs {e_ace:.2e}, ACE target agreement {agree:.6f}; "
          f"recall RX {r_rx:.6f} of {n_rx}, ACE {r_ace:.6f} of {n_ace}")
    assert e_rx <= BAR and e_ace <= BAR
    assert r_rx >= 0.999 and r_ace >= 0.999
    assert n_rx > 100 and n_ace > 100
