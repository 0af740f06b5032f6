"""Property tests of the device detectors: the reference's own detector tests
restated on the HIP path, through the C ABI. They hold for any correct
implementation, independently of the oracle and of its fixtures:
  - RX ranks planted anomalies above the whole background, and averages to the
    band count over the background - the chi-squared expectation, which
    catches a whitener off by a scale factor that a ranking cannot see
    (tests/test_detectors.cpp:82-133);
  - RX equals an explicit (x - mu)^T S^-1 (x - mu), and ACE its definition,
    from S^-1 formed in float64 by numpy (tests/test_detector.cpp:89-225);
  - ACE lies in [0, 1], peaks on its target, and its RX side output equals a
    standalone RX (tests/test_detectors.cpp:135-185);
  - CEM has unit gain on its target (w . d = 1) and suppresses the background
    (tests/test_detectors.cpp:187-212, tests/test_detector.cpp:227-249);
  - end to end, a generated frame's SAM detections land on planted pixels of
    the right target, find every target, and leave the background well clear
    of the threshold (tests/test_sources.cpp:83-144);
  - SAM: a target's own spectrum, at any gain, scores an angle of rounding
    size on it (the same bits at any power-of-two gain); a pixel's angle does
    not change with its gain; disjoint spectra are pi/2 apart
    (tests/test_sam.cpp:23-98);
  - the device whitener W = L^-1 whitens the covariance, W S W^T = I
    (tests/test_linalg.cpp:44-127, the Cholesky reconstruction);
  - a library whose band count differs from the cube's is refused
    (tests/test_detector.cpp, band mismatch);
  - the mixed-form statistics refuse a cube past one 4 GiB buffer view
    (no counterpart: the limit of this build's mixed form).
RX and ACE run in the plain FP64 form, the FP32 form and the mixed (FP16
hi/lo) form, each against its own bar: 1e-4 relative for the first two, as
the reference's tests, and 1e-3 for the mixed form.

Scenes: a correlated background (a ramp, three smooth spectral factors with
random amplitudes, white noise) with a few planted pixels of a spectrum
unlike it (a falling ramp with a narrow bump)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FORMS = ["plain", "fp32", "mixed"]
BAR = {"plain": 1e-4, "fp32": 1e-4, "mixed": 1e-3}


@pytest.fixture(scope="module")
def H():
    import torch
    import hsi_mi355x
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return hsi_mi355x


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


@pytest.fixture()
def ctx(H):
    c = H.Context(0)
    yield c
    c.close()


def correlated_scene(h, w, bands, seed, planted=3):
    rng = np.random.default_rng(seed)
    p = h * w
    t = np.linspace(0.0, 1.0, bands)
    mean = 0.25 + 0.2 * t
    factors = np.stack([np.cos(np.pi * k * t) for k in (1, 2, 3)])
    amp = rng.normal(0.0, 1.0, (3, p)) * np.array([0.04, 0.025, 0.015])[:, None]
    cube = mean[:, None] + factors.T @ amp + rng.normal(0.0, 0.008, (bands, p))
    target = 0.7 - 0.5 * t + 0.05 * np.exp(-((t - 0.5) / 0.08) ** 2)
    where = rng.choice(p, planted, replace=False)
    cube[:, where] = target[:, None] + rng.normal(0.0, 0.002, (bands, planted))
    background = np.ones(p, bool)
    background[where] = False
    return cube.astype(np.float32), target.astype(np.float32), mean.astype(np.float32), where, background


def device_scene(H, ctx, cube, h, w, library, loading, form):
    b = cube.shape[0]
    dev = H.to_device_cube(cube, h, w)
    ctx.set_library(np.asarray(library, np.float32))
    ctx.stats_accumulate(dev, h, w, b, precision=H.MIXED if form == "mixed" else H.PLAIN)
    ctx.stats_build(loading)
    ctx.stats_check()
    return dev


def precision_of(H, form):
    return {"plain": H.PLAIN, "fp32": H.FP32, "mixed": H.MIXED}[form]


def run_rx(H, torch, ctx, dev, h, w, b, form):
    rx = torch.empty(h * w, dtype=torch.float32, device="cuda")
    ctx.score(H.Detector.RX, dev, h, w, b, rx, precision=precision_of(H, form))
    torch.cuda.synchronize()
    return rx.cpu().numpy().astype(np.float64)


def explicit_inverse(cube, loading):
    x = cube.astype(np.float64)
    mu = x.mean(axis=1)
    s = (x @ x.T) / x.shape[1] - np.outer(mu, mu) + loading * np.eye(x.shape[0])
    return x - mu[:, None], np.linalg.inv(s), mu


@pytest.mark.parametrize("form", FORMS)
def test_rx_ranks_planted_anomalies_above_the_background(H, torch, ctx, form):
    h, w, b = 40, 48, 16
    cube, target, _, where, background = correlated_scene(h, w, b, seed=1)
    dev = device_scene(H, ctx, cube, h, w, target[None, :], 1e-6, form)
    rx = run_rx(H, torch, ctx, dev, h, w, b, form)
    assert rx[where].min() > rx[background].max()


@pytest.mark.parametrize("form", FORMS)
def test_rx_background_mean_is_the_band_count(H, torch, ctx, form):
    h, w, b = 60, 60, 12
    cube, target, _, _, background = correlated_scene(h, w, b, seed=2)
    dev = device_scene(H, ctx, cube, h, w, target[None, :], 1e-9, form)
    mean = run_rx(H, torch, ctx, dev, h, w, b, form)[background].mean()
    # three planted pixels inflate the covariance a little: a loose band
    assert 0.80 * b < mean < 1.05 * b, mean


@pytest.mark.parametrize("form", FORMS)
def test_rx_and_ace_match_an_explicit_inverse(H, torch, ctx, form):
    h, w, b = 40, 48, 16
    cube, target, mean, _, _ = correlated_scene(h, w, b, seed=5)
    library = np.stack([target, mean + 0.02])
    loading = 1e-6
    dev = device_scene(H, ctx, cube, h, w, library, loading, form)
    d, s_inv, mu = explicit_inverse(cube, loading)
    want_rx = np.einsum("ip,ij,jp->p", d, s_inv, d)
    rx = run_rx(H, torch, ctx, dev, h, w, b, form)
    assert np.max(np.abs(rx - want_rx) / np.maximum(np.abs(want_rx), 1.0)) <= BAR[form]
    # ACE of target t: (s^T S^-1 x)^2 / ((s^T S^-1 s)(x^T S^-1 x)), s = d_t - mu, best over t
    s = library.astype(np.float64).T - mu[:, None]
    num = (s.T @ s_inv @ d) ** 2
    den = np.einsum("it,ij,jt->t", s, s_inv, s)[:, None] * want_rx[None, :]
    want_ace = (num / den).max(axis=0)
    ace = torch.empty(h * w, dtype=torch.float32, device="cuda")
    which = torch.empty(h * w, dtype=torch.int32, device="cuda")
    side = torch.empty(h * w, dtype=torch.float32, device="cuda")
    ctx.score(H.Detector.ACE, dev, h, w, b, ace, which, side, precision=precision_of(H, form))
    torch.cuda.synchronize()
    assert np.max(np.abs(ace.cpu().numpy().astype(np.float64) - want_ace)) <= BAR[form]


@pytest.mark.parametrize("form", FORMS)
def test_ace_is_bounded_peaks_on_its_target_and_shares_rx(H, torch, ctx, form):
    h, w, b = 40, 48, 14
    cube, target, mean, where, _ = correlated_scene(h, w, b, seed=3)
    # a decoy that looks like the background, for the best-match reduction to reject
    dev = device_scene(H, ctx, cube, h, w, np.stack([target, mean]), 1e-6, form)
    p = h * w
    ace = torch.empty(p, dtype=torch.float32, device="cuda")
    which = torch.empty(p, dtype=torch.int32, device="cuda")
    side = torch.empty(p, dtype=torch.float32, device="cuda")
    ctx.score(H.Detector.ACE, dev, h, w, b, ace, which, side, precision=precision_of(H, form))
    torch.cuda.synchronize()
    a, t = ace.cpu().numpy(), which.cpu().numpy()
    assert (a >= 0.0).all() and (a <= 1.0).all()
    assert (a[where] >= 0.9).all() and (t[where] == 0).all()
    rx_alone = run_rx(H, torch, ctx, dev, h, w, b, form)
    assert np.max(np.abs(side.cpu().numpy().astype(np.float64) - rx_alone)) < 1e-3


def test_cem_has_unit_gain_on_its_target(H, torch, ctx):
    h, w, b = 40, 48, 14
    cube, target, _, where, background = correlated_scene(h, w, b, seed=4)
    dev = device_scene(H, ctx, cube, h, w, target[None, :], 1e-6, "plain")
    p = h * w
    cem = torch.empty(p, dtype=torch.float32, device="cuda")
    which = torch.empty(p, dtype=torch.int32, device="cuda")
    ctx.score(H.Detector.CEM, dev, h, w, b, cem, which)
    torch.cuda.synchronize()
    c = cem.cpu().numpy()
    assert ((c[where] > 0.8) & (c[where] < 1.2)).all()
    assert c[background].max() < 0.8
    # the filter itself: w . d = 1
    st = ctx.stats_read(b, 1)
    gain = float(st["cem_weight"][0].astype(np.float64) @ target.astype(np.float64))
    assert abs(gain - 1.0) < 1e-4


def test_sam_detections_land_on_the_planted_targets(H, torch, ctx):
    """A generated 256 x 256 x 113 frame with 4 targets planted at 0.5% each
    (abundance 0.6-1 over the background mix): SAM at 0.05 rad with
    suppression radius 2 detects only planted pixels, each with its own
    target, and every target at least once; no background pixel comes within
    twice the threshold."""
    h = w = 256
    b = 113
    cfg = H.synth_config(h, w, b, 4)
    frame, truth = H.synth_frame(cfg, 0, truth=True, threads=8)
    library = H.synth_library(cfg)
    p = h * w
    cube = torch.from_numpy(frame).cuda()
    ctx.set_library(library)
    angle = torch.empty(p, dtype=torch.float32, device="cuda")
    which = torch.empty(p, dtype=torch.int32, device="cuda")
    ctx.score(H.Detector.SAM, cube, h, w, b, angle, which)
    cap = 4096
    out = torch.zeros((cap, 4), dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.detect(angle, which, h, w, H.DetectionParams(0.05, 2, cap, False), out, count)
    torch.cuda.synchronize()
    found = int(count.item())
    assert 0 < found <= cap
    dets = H.detections_to_numpy(out, found)
    pix = dets["y"].astype(np.int64) * w + dets["x"]
    assert (truth[pix] != 0).all()
    assert (truth[pix] == dets["target"] + 1).all()
    assert set(dets["target"].tolist()) == {0, 1, 2, 3}
    assert angle.cpu().numpy()[truth[:p] == 0].min() > 1.5 * 0.05


def test_sam_closed_forms(H, torch, ctx):
    rng = np.random.default_rng(9)
    b = 16
    library = rng.uniform(0.05, 1.0, (3, b)).astype(np.float32)
    library[2, 8:] = 0.0  # target 2 lives in bands 0-7 only
    base = rng.uniform(0.05, 1.0, b).astype(np.float32)
    disjoint = np.zeros(b, np.float32)
    disjoint[8:] = rng.uniform(0.1, 1.0, 8)
    pixels = [library[0], library[1], library[2], library[1] * np.float32(0.25), library[2] * np.float32(8.0),
              library[1] * np.float32(3.7)]
    gains = [1.0, 0.01, 0.5, 3.0, 250.0, 0.125, 64.0]
    pixels += [base * np.float32(g) for g in gains]
    pixels += [disjoint]
    cube = np.stack(pixels, axis=1)  # [bands, pixels]
    h, w = 1, cube.shape[1]
    dev = H.to_device_cube(cube, h, w)
    ctx.set_library(library)
    angle = torch.empty(w, dtype=torch.float32, device="cuda")
    which = torch.empty(w, dtype=torch.int32, device="cuda")
    ctx.score(H.Detector.SAM, dev, h, w, b, angle, which)
    torch.cuda.synchronize()
    a, t = angle.cpu().numpy(), which.cpu().numpy()
    # a target's own spectrum: the angle is at most the acos of a cosine one
    # float step from 1 (the target norms are floats, as the reference's)
    assert (a[:3] <= 5e-4).all() and list(t[:3]) == [0, 1, 2]
    assert a[3] == a[1] and a[4] == a[2] and t[3] == 1 and t[4] == 2
    assert a[5] <= 5e-4 and t[5] == 1
    scaled = a[6:6 + len(gains)]
    assert np.abs(scaled.astype(np.float64) - scaled[0]).max() <= 1e-6
    exact = [i for i, g in enumerate(gains) if np.log2(g) == np.round(np.log2(g))]
    assert (scaled[exact] == scaled[0]).all()  # a power-of-two gain changes no bit
    # the disjoint pixel is orthogonal to target 2 and not to the others
    ctx.set_library(library[2:3])
    ctx.score(H.Detector.SAM, dev, h, w, b, angle, which)
    torch.cuda.synchronize()
    assert angle.cpu().numpy()[-1] == np.float32(np.pi / 2)


@pytest.mark.parametrize("bands", [16, 113])
def test_whitener_whitens_the_covariance(H, torch, ctx, bands):
    h, w = 40, 48
    cube, target, _, _, _ = correlated_scene(h, w, bands, seed=6)
    loading = 1e-6
    device_scene(H, ctx, cube, h, w, target[None, :], loading, "plain")
    whitener = ctx.stats_read(bands, 1)["whitener"].astype(np.float64)
    x = cube.astype(np.float64)
    mu = x.mean(axis=1)
    s = (x @ x.T) / x.shape[1] - np.outer(mu, mu) + loading * np.eye(bands)
    assert np.allclose(np.triu(whitener, 1), 0.0)  # lower triangular
    assert np.abs(whitener @ s @ whitener.T - np.eye(bands)).max() <= 1e-5


def test_library_band_mismatch_is_refused(H, torch, ctx):
    h, w, b = 8, 8, 12
    cube, target, _, _, _ = correlated_scene(h, w, b, seed=7)
    dev = H.to_device_cube(cube, h, w)
    ctx.set_library(np.ones((2, b + 1), np.float32))
    out = torch.empty(h * w, dtype=torch.float32, device="cuda")
    idx = torch.empty(h * w, dtype=torch.int32, device="cuda")
    with pytest.raises(Exception) as info:
        ctx.score(H.Detector.SAM, dev, h, w, b, out, idx)
    assert "bands" in str(info.value)
    ctx.stats_accumulate(dev, h, w, b)
    with pytest.raises(Exception) as info:
        ctx.stats_build(1e-6)
        ctx.stats_check()
    assert "bands" in str(info.value)


@pytest.mark.gpu
def test_mixed_statistics_refuse_a_cube_past_one_buffer_view(H, torch, ctx):
    """The mixed-form statistics read the cube through one buffer view: a
    plane stride that puts the cube past 4 GiB, or only the staging offsets of
    its padding rows past 2^32 (209 FP16 bands staged as 224 rows), is refused
    with an argument error before anything is launched (the cube itself is a
    few bytes here and is never read)."""
    import ctypes as C
    lib = H.lib()
    small = torch.zeros(64, dtype=torch.float32, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for bands, stride, dtype in [(4, 1 << 30, H.F16), (209, 10_000_000, H.F16), (4, 1 << 28, H.F32)]:
        st = lib.hsi_stats_accumulate(ctx.handle, C.c_void_p(small.data_ptr()), dtype, 2, 2, bands,
                                      C.c_size_t(stride), H.MIXED, None, stream)
        assert st == H.ERR_INVALID_ARG, (bands, stride, H.status_name(st))
        assert "4 GiB" in lib.hsi_ctx_last_error(ctx.handle).decode()
