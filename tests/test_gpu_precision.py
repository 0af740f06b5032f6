"""The precision form end to end against the FP64 oracle, on a scene large
enough for the covariance to be well estimated (256x256 px x 224 bands, 16
targets - the band count and library of the 224-band precision workload).

mixed: FP16 hi/lo moments with FP32 tiles reduced in FP64, FP64 statistics
build, FP16 MFMA whitening of an FP16 cube. The bar is 1e-3: RX relative
(floor 1), ACE absolute. The oracle scores the same FP16 values the device
reads, with statistics from the exact moments of those values, so the error
is the arithmetic's alone. The plain path on the same cube must stay within
1e-5, and so must the FP32 form (whitening on FP32 matrix cores, FP32
accumulation) on the plain statistics.
"""
import numpy as np
import pytest

from gpu_util import rel_err

pytestmark = pytest.mark.gpu

MIXED_BAR = 1e-3
PLAIN_BAR = 1e-5
LOADING = 1e-6


@pytest.fixture(scope="module")
def scene(oracle_lib):
    import hsi_mi355x as H
    h, w, b, t = 256, 256, 224, 16
    cfg = H.synth_config(h, w, b, t, seed=424242)
    lib = H.synth_library(cfg)
    host = H.synth_frame(cfg, 0, threads=8)
    halves = host.astype(np.float16).astype(np.float32)  # the values an FP16 cube holds
    count, s1, s2 = oracle_lib.moments(halves, pixels=h * w, threads=8)
    st = oracle_lib.build_statistics(count, s1, s2, lib, LOADING)
    assert st.valid
    want_ace, want_t, want_rx = oracle_lib.ace_best(halves, st, pixels=h * w, threads=8)
    return dict(H=H, h=h, w=w, b=b, lib=lib, halves=halves, ace=want_ace, target=want_t, rx=want_rx)


@pytest.mark.parametrize("form", ["plain", "mixed", "fp32"])
def test_rx_ace_path_against_fp64_oracle(scene, form):
    import torch
    H = scene["H"]
    h, w, b = scene["h"], scene["w"], scene["b"]
    p = h * w
    print(f"precision form {form}: {h}x{w}x{b}", flush=True)
    c32 = torch.from_numpy(scene["halves"]).cuda()
    c16 = c32.half()
    ctx = H.Context(0)
    ctx.set_library(scene["lib"])
    prec = {"plain": H.PLAIN, "mixed": H.MIXED, "fp32": H.FP32}[form]
    ace = torch.empty(p, dtype=torch.float32, device="cuda")
    rx = torch.empty(p, dtype=torch.float32, device="cuda")
    tg = torch.empty(p, dtype=torch.int32, device="cuda")
    # the FP32 form scores on the plain (FP64) statistics
    ctx.stats_accumulate(c32, h, w, b, precision=H.MIXED if form == "mixed" else H.PLAIN)
    ctx.stats_build(LOADING)
    ctx.stats_check()
    ctx.score(H.Detector.ACE, c16, h, w, b, ace, tg, rx, precision=prec)
    torch.cuda.synchronize()
    got_ace, got_rx, got_t = ace.cpu().numpy(), rx.cpu().numpy(), tg.cpu().numpy()
    ctx.close()
    e_rx = rel_err(got_rx, scene["rx"], 1.0).max()
    e_ace = np.abs(got_ace.astype(np.float64) - scene["ace"]).max()
    agree = (got_t == scene["target"]).mean()
    print(f"{form}: RX max rel {e_rx:.3g}, ACE max abs {e_ace:.3g}, ACE target agreement {agree:.6f}")
    bar = MIXED_BAR if form == "mixed" else PLAIN_BAR
    assert e_rx <= bar and e_ace <= bar
    assert agree >= 0.999
