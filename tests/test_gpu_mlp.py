"""The learned pre-screen on the GPU against its CPU oracle
(oracle/mlp_ref.py, itself pinned by finite differences): gradients of the
device forward/backward kernels, one Adam step, bf16 inference, and the exact
rescoring of candidate pixels against the full-frame plain path."""
import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu

BANDS = 113


@pytest.fixture(scope="module")
def H():
    import torch
    import hsi_mi355x
    assert torch.cuda.is_available()
    return hsi_mi355x


def batch(n=3000, bands=BANDS, seed=5, pos=0.1):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 0.3, (bands, n)).astype(np.float32)
    y = (rng.random(n) < pos).astype(np.float32)
    return x, y


def test_gradients_match_oracle(H):
    import torch
    ctx = H.Context(0)
    ctx.mlp_init(BANDS, seed=9)
    w = ctx.mlp_get_weights().astype(np.float64)
    x, y = batch()
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    g = torch.zeros(H.mlp_param_count(BANDS), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    print("mlp gradients, 3000 samples x 113 bands", flush=True)
    ctx.mlp_gradients(xd, yd, 4.0, g, loss)
    torch.cuda.synchronize()
    want_loss, want = mlp_ref.loss_and_grad(w, x.astype(np.float64), y, 4.0, BANDS)
    got = g.cpu().numpy().astype(np.float64)
    scale = np.abs(want).max()
    err = np.abs(got - want).max() / scale
    print(f"gradient max error {err:.3g} of max |g| {scale:.3g}; loss {loss.item():.6f} vs {want_loss:.6f}")
    assert err <= 1e-4
    assert abs(loss.item() - want_loss) <= 1e-5 * max(1.0, want_loss)
    # every parameter block is exercised
    offs = np.cumsum([0, 64 * BANDS, 64, 32 * 64, 32, 32, 1])
    for a, b in zip(offs[:-1], offs[1:]):
        assert np.abs(want[a:b]).max() > 0
        assert np.abs(got[a:b] - want[a:b]).max() <= 1e-4 * scale
    ctx.close()


def test_adam_step_matches_oracle(H):
    import torch
    ctx = H.Context(0)
    ctx.mlp_init(BANDS, seed=4)
    w0 = ctx.mlp_get_weights().astype(np.float64)
    x, y = batch(n=2048, seed=6)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    p = H.MlpTrainParams(learning_rate=1e-3, pos_weight=2.0)
    ctx.mlp_train_step(xd, yd, p)
    ctx.mlp_train_step(xd, yd, p)
    torch.cuda.synchronize()
    got = ctx.mlp_get_weights().astype(np.float64)
    w, m, v = w0, np.zeros_like(w0), np.zeros_like(w0)
    for t in (1, 2):
        _, g = mlp_ref.loss_and_grad(w, x.astype(np.float64), y, 2.0, BANDS)
        w, m, v = mlp_ref.adam_step(w, m, v, g, t, 1e-3, 0.9, 0.999, 1e-8)
    # Adam normalises each step to ~lr, so a gradient within rounding of zero
    # may move either way; elsewhere the update must agree closely.
    _, g1 = mlp_ref.loss_and_grad(w0, x.astype(np.float64), y, 2.0, BANDS)
    solid = np.abs(g1) > 1e-5 * np.abs(g1).max()
    err = np.abs(got - w)
    print(f"adam: max error {err[solid].max():.3g} on {solid.sum()} solid entries, "
          f"{err.max():.3g} overall")
    assert err[solid].max() <= 2e-6
    assert err.max() <= 4.1e-3
    ctx.close()


def test_training_reduces_loss(H):
    import torch
    ctx = H.Context(0)
    ctx.mlp_init(BANDS, seed=1)
    rng = np.random.default_rng(2)
    n = 4096
    x = rng.normal(0, 0.3, (BANDS, n)).astype(np.float32)
    direction = rng.normal(0, 1, BANDS).astype(np.float32)
    y = ((direction @ x) > 1.0).astype(np.float32)  # separable by a hyperplane
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    p = H.MlpTrainParams(learning_rate=3e-3, pos_weight=1.0)
    losses = []
    for step in range(300):
        ctx.mlp_train_step(xd, yd, p, loss)
        if step % 50 == 0 or step == 299:
            torch.cuda.synchronize()
            losses.append(loss.item())
    print("training loss", [round(v, 4) for v in losses])
    assert losses[-1] < 0.25 * losses[0]
    ctx.close()


def test_inference_matches_bf16_oracle(H):
    import torch
    ctx = H.Context(0)
    ctx.mlp_init(BANDS, seed=12)
    w = ctx.mlp_get_weights().astype(np.float64)
    h, wd = 13, 29
    rng = np.random.default_rng(8)
    cube = rng.uniform(0.0, 0.9, (BANDS, h * wd)).astype(np.float32)
    dev = H.to_device_cube(cube, h, wd)
    logit = torch.full((h * wd,), float("nan"), dtype=torch.float32, device="cuda")
    print("mlp inference 13x29x113", flush=True)
    ctx.mlp_infer(dev, h, wd, BANDS, logit, center=False)
    torch.cuda.synchronize()
    got = logit.cpu().numpy().astype(np.float64)
    want = mlp_ref.forward_bf16(w, cube.astype(np.float64), BANDS)
    exact = mlp_ref.forward(w, cube.astype(np.float64), BANDS)[0]
    err = np.abs(got - want)
    print(f"inference: max |logit - bf16 oracle| {err.max():.3g}, median {np.median(err):.3g}; "
          f"vs float64 network {np.abs(got - exact).max():.3g}")
    assert np.isfinite(got).all()
    assert np.median(err) <= 1e-4 * (1 + np.abs(want).max())
    assert err.max() <= 1e-2 * (1 + np.abs(want).max())
    # the fp16 cube path reads the same values when they are halves already
    halves = cube.astype(np.float16)
    dev16 = H.to_device_cube(halves.astype(np.float32), h, wd, dtype="f16")
    ctx.mlp_infer(dev16, h, wd, BANDS, logit, center=False)
    torch.cuda.synchronize()
    dev32 = H.to_device_cube(halves.astype(np.float32), h, wd)
    l2 = torch.empty_like(logit)
    ctx.mlp_infer(dev32, h, wd, BANDS, l2, center=False)
    torch.cuda.synchronize()
    assert torch.equal(logit, l2)
    ctx.close()


@pytest.mark.parametrize("det", ["rx", "ace"])
def test_candidate_rescoring_equals_full_frame(H, det):
    import torch
    from conftest import golden
    g = golden("scene_32x40x113_t4")
    b, h, w = g["cube"].shape
    p = h * w
    cube = H.to_device_cube(g["cube"], h, w)
    ctx = H.Context(0)
    ctx.set_library(g["library"])
    ctx.stats_accumulate(cube, h, w, b)
    ctx.stats_build(1e-6)
    D = H.Detector.RX if det == "rx" else H.Detector.ACE
    f32 = dict(dtype=torch.float32, device="cuda")
    full_s, full_a = torch.empty(p, **f32), torch.empty(p, **f32)
    full_t = torch.empty(p, dtype=torch.int32, device="cuda")
    ctx.score(D, cube, h, w, b, full_s, full_t if det == "ace" else None, full_a if det == "ace" else None)
    # candidates: every third pixel, as raster-order detection records
    idx = np.arange(0, p, 3)
    rec = np.zeros((idx.size + 5, 4), np.int32)
    rec[:idx.size, 0] = idx % w
    rec[:idx.size, 1] = idx // w
    cand = torch.from_numpy(rec).cuda()
    count = torch.tensor([idx.size], dtype=torch.int32, device="cuda")
    s = torch.full((p,), float("nan"), **f32)
    a = torch.full((p,), float("nan"), **f32)
    t = torch.full((p,), -7, dtype=torch.int32, device="cuda")
    print(f"candidate rescoring {det}: {idx.size} of {p}", flush=True)
    ctx.score_candidates(D, cube, h, w, b, cand, count, s, t if det == "ace" else None,
                         a if det == "ace" else None)
    torch.cuda.synchronize()
    got, want = s.cpu().numpy(), full_s.cpu().numpy()
    mask = np.zeros(p, bool)
    mask[idx] = True
    assert np.array_equal(got[mask].view(np.uint32), want[mask].view(np.uint32))
    assert np.all(got[~mask] == 0)
    if det == "ace":
        assert np.array_equal(t.cpu().numpy()[mask], full_t.cpu().numpy()[mask])
        assert np.all(t.cpu().numpy()[~mask] == -1)
        assert np.array_equal(a.cpu().numpy()[mask], full_a.cpu().numpy()[mask])
    # a list that overflowed its capacity scores the whole frame
    count.fill_(10 ** 6)
    ctx.score_candidates(D, cube, h, w, b, cand[:64], count, s)
    torch.cuda.synchronize()
    assert np.array_equal(s.cpu().numpy().view(np.uint32), want.view(np.uint32))
    ctx.close()


@pytest.mark.parametrize("name", ["scene_32x40x113_t4", "scene_13x37x113_t8"])
def test_one_pass_screen_matches_logits_and_full_frame(H, name):
    """hsi_mlp_screen's candidates are exactly the pixels whose logit (as
    hsi_mlp_infer computes it) reaches the cut, with those logits and their
    cube spectra in the distinct rows their records name, listed the same way
    on a second run; rescoring from those rows equals the full-frame plain
    ACE bit for bit; an
    overflowing list falls back to the whole frame."""
    import torch
    from conftest import golden
    g = golden(name)
    b, h, w = g["cube"].shape
    p = h * w
    cube = H.to_device_cube(g["cube"], h, w)
    ctx = H.Context(0)
    ctx.set_library(g["library"])
    ctx.stats_accumulate(cube, h, w, b)
    ctx.stats_build(1e-6)
    ctx.mlp_init(b, seed=21)
    f32 = dict(dtype=torch.float32, device="cuda")
    i32 = dict(dtype=torch.int32, device="cuda")
    logit = torch.empty(p, **f32)
    ctx.mlp_infer(cube, h, w, b, logit, center=True)
    torch.cuda.synchronize()
    lg = logit.cpu().numpy()
    cut = float(np.quantile(lg, 0.8))
    want = np.flatnonzero(lg >= cut)
    cap = p + 3
    cand = torch.full((cap, 4), -9, **i32)
    count = torch.full((1,), 12345, **i32)
    spectra = torch.full((cap, (b + 31) // 32 * 32), float("nan"), **f32)
    logit2 = torch.empty(p, **f32)
    print(f"screen {name}: cut {cut:.4f}, {want.size} of {p} expected", flush=True)
    ctx.mlp_screen(cube, h, w, b, cut, cand, count, spectra, logit2)
    torch.cuda.synchronize()
    n = int(count.item())
    assert n == want.size
    assert torch.equal(logit, logit2)
    rec = cand[:n].cpu().numpy()
    got_px = rec[:, 1].astype(np.int64) * w + rec[:, 0]
    assert np.array_equal(np.sort(got_px), want)
    rows = rec[:, 2]
    assert np.all((rows >= 0) & (rows < cap)) and np.unique(rows).size == n
    assert np.array_equal(rec[:, 3].view(np.float32), lg[got_px])
    host = g["cube"].reshape(b, p)
    again = torch.full((cap, 4), -9, **i32)
    ctx.mlp_screen(cube, h, w, b, cut, again, count, spectra)
    torch.cuda.synchronize()
    assert torch.equal(again[:n], cand[:n])
    sp = spectra.cpu().numpy()
    assert np.array_equal(sp[rows, :b], host[:, got_px].T)
    assert np.all(sp[rows, b:] == 0)
    # without spectra the records carry row 0
    ctx.mlp_screen(cube, h, w, b, cut, again, count)
    torch.cuda.synchronize()
    assert np.array_equal(again[:n].cpu().numpy()[:, [0, 1, 3]], rec[:, [0, 1, 3]])
    assert np.all(again[:n].cpu().numpy()[:, 2] == 0)

    full_s, full_a = torch.empty(p, **f32), torch.empty(p, **f32)
    full_t = torch.empty(p, **i32)
    ctx.score(H.Detector.ACE, cube, h, w, b, full_s, full_t, full_a)
    s = torch.full((p,), float("nan"), **f32)
    a = torch.full((p,), float("nan"), **f32)
    t = torch.full((p,), -7, **i32)
    ctx.score_candidate_spectra(H.Detector.ACE, cube, h, w, b, cand, count, spectra, s, t, a)
    torch.cuda.synchronize()
    mask = np.zeros(p, bool)
    mask[want] = True
    for got, ref, fill in ((s, full_s, 0.0), (a, full_a, 0.0), (t, full_t, -1)):
        gv, rv = got.cpu().numpy(), ref.cpu().numpy()
        assert np.array_equal(gv[mask].view(np.uint32), rv[mask].view(np.uint32))
        assert np.all(gv[~mask] == fill)

    # capacity below the count: the list overflows, the rescoring covers the frame
    small = want.size // 2
    cand_s = torch.full((small, 4), -9, **i32)
    spec_s = torch.empty((small, (b + 31) // 32 * 32), **f32)
    ctx.mlp_screen(cube, h, w, b, cut, cand_s, count, spec_s)
    ctx.score_candidate_spectra(H.Detector.ACE, cube, h, w, b, cand_s, count, spec_s, s, t, a)
    torch.cuda.synchronize()
    assert int(count.item()) > small
    assert np.array_equal(s.cpu().numpy().view(np.uint32), full_s.cpu().numpy().view(np.uint32))
    assert np.array_equal(t.cpu().numpy(), full_t.cpu().numpy())
    ctx.close()


@pytest.mark.parametrize("det", ["rx", "ace"])
@pytest.mark.parametrize("name", ["scene_32x40x113_t4", "scene_13x37x113_t8", "scene_20x20x224_t16"])
def test_mixed_rescoring_of_candidate_rows_matches_full_frame_mixed(H, name, det):
    """The precision form over candidate spectra rows (any row order, a pitch
    past the band count whose extra elements are NaN, records outside the
    frame or naming no row skipped) gives each listed pixel the full-frame
    precision-form score (FP32 rounding of the accumulation apart: 2e-6
    relative) and the same ACE target; every other pixel keeps the fill; an
    overflowing list falls back to the whole frame in that form."""
    import torch
    from conftest import golden
    g = golden(name)
    b, h, w = g["cube"].shape
    p = h * w
    D = H.Detector.RX if det == "rx" else H.Detector.ACE
    cube = H.to_device_cube(g["cube"], h, w)
    ctx = H.Context(0)
    ctx.set_library(g["library"])
    ctx.stats_accumulate(cube, h, w, b)
    ctx.stats_build(1e-6)
    f32 = dict(dtype=torch.float32, device="cuda")
    i32 = dict(dtype=torch.int32, device="cuda")
    full_s, full_a = torch.empty(p, **f32), torch.empty(p, **f32)
    full_t = torch.empty(p, **i32)
    ctx.score(D, cube, h, w, b, full_s, full_t, full_a, precision=H.MIXED)
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(p, size=max(1, p // 3), replace=False))
    n = pick.size
    cap = n + 7
    pitch = (b + 3) // 4 * 4 + 4
    rows = rng.permutation(cap)[:n]
    host = g["cube"].reshape(b, p)
    spec = np.full((cap, pitch), np.nan, np.float32)
    spec[rows, :b] = host[:, pick].T
    rec = np.zeros((cap, 4), np.int32)
    rec[:n, 0] = pick % w
    rec[:n, 1] = pick // w
    rec[:n, 2] = rows
    # two extra records that name no pixel: outside the frame, and a row past the capacity
    rec[n] = (w, 0, rows[0], 0)
    rec[n + 1] = (0, 0, cap, 0)
    cand = torch.from_numpy(rec).cuda()
    spectra = torch.from_numpy(spec).cuda()
    count = torch.tensor([n + 2], **i32)
    s = torch.full((p,), float("nan"), **f32)
    a = torch.full((p,), float("nan"), **f32)
    t = torch.full((p,), -7, **i32)
    ctx.score_candidate_spectra(D, cube, h, w, b, cand, count, spectra, s, t, a, precision=H.MIXED)
    torch.cuda.synchronize()
    mask = np.zeros(p, bool)
    mask[pick] = True
    fs = full_s.cpu().numpy()
    gs = s.cpu().numpy()
    rel = np.abs(gs[mask] - fs[mask]) / np.maximum(np.abs(fs[mask]), 1e-30)
    print(f"{name} {det}: {n} rows, max relative difference {rel.max():.3g}", flush=True)
    assert rel.max() <= 2e-6
    assert np.all(gs[~mask] == 0.0)
    if det == "ace":
        assert np.array_equal(t.cpu().numpy()[mask], full_t.cpu().numpy()[mask])
        assert np.all(t.cpu().numpy()[~mask] == -1)
        ga, fa = a.cpu().numpy(), full_a.cpu().numpy()
        assert np.all(np.abs(ga[mask] - fa[mask]) <= 2e-6 * np.abs(fa[mask]))
    # a count above the capacity: the whole frame in the precision form
    count.fill_(cap + 1)
    ctx.score_candidate_spectra(D, cube, h, w, b, cand, count, spectra, s, t, a, precision=H.MIXED)
    torch.cuda.synchronize()
    gs = s.cpu().numpy()
    assert np.all(np.abs(gs - fs) <= 2e-6 * np.abs(fs))
    if det == "ace":
        assert np.array_equal(t.cpu().numpy(), full_t.cpu().numpy())
    # rows need 16-byte alignment in this form
    with pytest.raises(H.HsiError):
        ctx.score_candidate_spectra(D, cube, h, w, b, cand, count, spectra[:, 1:], s, t, a, precision=H.MIXED)
    ctx.close()
