"""GPU parity against the golden fixtures (written by the compiled reference)
and the CPU oracle. Every call goes through the C ABI.

Bars (the reference-scorer parity the path is held to):
  - SAM angles within 1e-6 rad of the oracle; winning target identical
    unless the two best angles are within 1e-6 rad of each other.
  - CEM given the reference's own statistics: bit-identical (the kernel forms
    the same exact double products and sums them in the same band order).
  - RX / ACE scores within 1e-5 relative (floor 1 for the unit-scaled ACE);
    ACE target identical unless the two best scores are within that band.
  - Moments within 1e-12 relative (exact products, different sum order);
    statistics built on the device within a few float ulps.
  - Detection lists from the same score map: identical, element by element.
"""
import numpy as np
import pytest

from conftest import SCENES, golden
from gpu_util import (ANGLE_TOL, FLOOR, REL, device_buffers, near_cut, rel_err, same_bits,
                      upload_golden_stats)

pytestmark = pytest.mark.gpu


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


def load_scene(H, torch, g, dtype="f32"):
    b, h, w = g["cube"].shape
    return H.to_device_cube(g["cube"], h, w, dtype=dtype), b, h, w


def sam_bars(got_s, got_t, g, pixels):
    want_s, want_t = g["sam_score"], g["sam_target"]
    err = np.abs(got_s.astype(np.float64) - want_s.astype(np.float64))
    assert err.max() <= ANGLE_TOL, f"max angle error {err.max()}"
    diff = np.nonzero(got_t != want_t)[0]
    if diff.size:
        allv = g["sam_all"]
        for p in diff:
            gap = abs(float(allv[got_t[p], p]) - float(want_s[p]))
            assert gap <= ANGLE_TOL, f"pixel {p}: target {got_t[p]} vs {want_t[p]}, gap {gap}"
    return float(err.max()), int((got_s.view(np.uint32) == want_s.view(np.uint32)).sum())


@pytest.mark.parametrize("name", SCENES)
def test_sam_matches_reference(H, torch, ctx, name):
    g = golden(name)
    cube, b, h, w = load_scene(H, torch, g)
    p = h * w
    ctx.set_library(g["library"], g["norms"])
    buf = device_buffers(torch, p)
    ctx.score(H.Detector.SAM, cube, h, w, b, buf["score"], buf["target"])
    torch.cuda.synchronize()
    s = buf["score"].cpu().numpy()
    t = buf["target"].cpu().numpy()
    maxerr, exact = sam_bars(s, t, g, p)
    print(f"{name}: SAM max |d angle| {maxerr:.3g} rad, bit-identical {exact}/{p}")


@pytest.mark.parametrize("name", SCENES)
def test_sam_fp16_cube_matches_oracle_on_the_same_halves(H, torch, ctx, oracle_lib, name):
    g = golden(name)
    b, h, w = g["cube"].shape
    halves = g["cube"].astype(np.float16).astype(np.float32)
    cube = H.to_device_cube(halves, h, w, dtype="f16")
    ctx.set_library(g["library"], g["norms"])
    buf = device_buffers(torch, h * w)
    ctx.score(H.Detector.SAM, cube, h, w, b, buf["score"], buf["target"])
    torch.cuda.synchronize()
    want_s, want_t = oracle_lib.sam_best(halves, g["library"], g["norms"])
    got = buf["score"].cpu().numpy()
    assert np.abs(got.astype(np.float64) - want_s).max() <= ANGLE_TOL
    assert (buf["target"].cpu().numpy() == want_t).mean() > 0.999


def test_sam_full_frame_against_oracle(H, torch, ctx, oracle_lib):
    """A 1021 x 1031 x 113 frame of the seeded generator with 8 targets
    (1,052,651 pixels: more than one sweep of the SAM grid, and a ragged last
    four-pixel group) against the oracle's sam_best; the fused SAM+CEM launch
    equals the separate ones bit for bit at this size."""
    h, w, b, t = 1021, 1031, 113, 8
    cfg = H.synth_config(h, w, b, t, seed=20261203 + 77)
    host = H.synth_frame(cfg, 0, threads=8)
    lib = np.asarray(H.synth_library(cfg), np.float32)
    p = h * w
    cube = torch.from_numpy(host).cuda()
    ctx.set_library(lib)
    sam = device_buffers(torch, p)
    ctx.score(H.Detector.SAM, cube, h, w, b, sam["score"], sam["target"])
    torch.cuda.synchronize()
    got_s, got_t = sam["score"].cpu().numpy(), sam["target"].cpu().numpy()
    want_s, want_t = oracle_lib.sam_best(host, lib, pixels=p, threads=8)
    err = np.abs(got_s.astype(np.float64) - want_s.astype(np.float64))
    assert err.max() <= ANGLE_TOL, f"max angle error {err.max()}"
    diff = np.nonzero(got_t != want_t)[0]
    if diff.size:
        # a different winner only where its angle ties the oracle's best
        allv = oracle_lib.sam_all(np.ascontiguousarray(host[:, diff]), lib)
        gap = np.abs(allv[got_t[diff], np.arange(diff.size)].astype(np.float64) - want_s[diff])
        assert gap.max() <= ANGLE_TOL
    exact = int((got_s.view(np.uint32) == want_s.view(np.uint32)).sum())
    print(f"full frame: SAM max |d angle| {err.max():.3g} rad, bit-identical {exact}/{p}, "
          f"different winners {diff.size}")
    assert exact >= 0.999 * p
    ctx.stats_accumulate(cube, h, w, b)
    ctx.stats_build(1e-6)
    ctx.stats_check()
    fused_s, fused_c = device_buffers(torch, p), device_buffers(torch, p)
    ctx.score_sam_cem(cube, h, w, b, fused_s["score"], fused_s["target"], fused_c["score"], fused_c["target"])
    cem = device_buffers(torch, p)
    ctx.score(H.Detector.CEM, cube, h, w, b, cem["score"], cem["target"])
    torch.cuda.synchronize()
    for x, y in ((fused_s, sam), (fused_c, cem)):
        assert same_bits(x["score"].cpu().numpy(), y["score"].cpu().numpy())
        assert np.array_equal(x["target"].cpu().numpy(), y["target"].cpu().numpy())


def test_sam_edge_cases(H, torch, ctx):
    g = golden("edge_cases")
    cube = H.to_device_cube(g["edge_cube"], 1, 8)
    ctx.set_library(g["edge_library"], g["edge_norms"])
    buf = device_buffers(torch, 8)
    ctx.score(H.Detector.SAM, cube, 1, 8, 6, buf["score"], buf["target"])
    torch.cuda.synchronize()
    s = buf["score"].cpu().numpy()
    t = buf["target"].cpu().numpy()
    want_s, want_t = g["edge_sam_score"], g["edge_sam_target"]
    assert np.abs(s.astype(np.float64) - want_s).max() <= ANGLE_TOL
    assert np.array_equal(t, want_t)  # zero pixel -> target 0, duplicate -> first, zero-norm target skipped
    assert not np.isnan(s).any()
    assert s[0] == np.float32(np.pi / 2)


def near_tie_scene(bands, targets, seed):
    """Library and pixels that crowd the SAM winner: a target repeated at
    other scales (same direction, cosines a rounding apart), one nudged by a
    few ulps, pixels equal to or scaled from targets (cosine clamps at 1),
    negated targets (cosine -1), tiny pixels, and random mixtures."""
    rng = np.random.default_rng(seed)
    lib = rng.uniform(0.05, 0.9, (targets, bands)).astype(np.float32)
    lib[1] = lib[0] * np.float32(3.0)
    lib[2] = lib[0] * np.float32(0.7)
    lib[3] = np.nextafter(lib[0], np.float32(2.0))
    pix = []
    for t in range(targets):
        pix += [lib[t], lib[t] * np.float32(0.37), -lib[t], lib[t] * np.float32(1e-18)]
    pix += list(rng.dirichlet(np.ones(targets), 200).astype(np.float32) @ lib)
    pix += list(lib[0] + rng.normal(0, 1e-6, (40, bands)).astype(np.float32))
    cube = np.stack(pix).astype(np.float32)  # [pixels, bands]
    return np.ascontiguousarray(cube.T), lib


@pytest.mark.parametrize("targets", [4, 8, 16])
def test_sam_ranked_winner_on_near_ties(H, torch, ctx, oracle_lib, targets):
    """Targets are ranked with reciprocal products and only the contenders take
    the exact quotient; the winner and angle must still be the reference's."""
    bands = 113
    cube2, lib = near_tie_scene(bands, targets, 40 + targets)
    p = cube2.shape[1]
    norms = oracle_lib.library_norms(lib)
    cube = H.to_device_cube(cube2.reshape(bands, 1, p), 1, p)
    ctx.set_library(lib, norms)
    buf = device_buffers(torch, p)
    ctx.score(H.Detector.SAM, cube, 1, p, bands, buf["score"], buf["target"])
    torch.cuda.synchronize()
    got_s = buf["score"].cpu().numpy()
    got_t = buf["target"].cpu().numpy()
    want_s, want_t = oracle_lib.sam_best(cube2, lib, norms)
    allv = oracle_lib.sam_all(cube2, lib, norms)
    assert np.abs(got_s.astype(np.float64) - want_s).max() <= ANGLE_TOL
    for q in np.nonzero(got_t != want_t)[0]:
        gap = abs(float(allv[got_t[q], q]) - float(want_s[q]))
        assert gap <= ANGLE_TOL, f"pixel {q}: target {got_t[q]} vs {want_t[q]}"
    same = int((got_t == want_t).sum())
    print(f"near ties, {targets} targets: winner identical {same}/{p}, "
          f"bit-identical angles {int((got_s.view(np.uint32) == want_s.view(np.uint32)).sum())}/{p}")
    assert same >= p - 2


@pytest.mark.parametrize("name", SCENES)
def test_cem_with_reference_statistics_is_bit_identical(H, torch, ctx, name):
    g = golden(name)
    cube, b, h, w = load_scene(H, torch, g)
    ctx.set_library(g["library"], g["norms"])
    upload_golden_stats(ctx, g)
    buf = device_buffers(torch, h * w)
    ctx.score(H.Detector.CEM, cube, h, w, b, buf["score"], buf["target"])
    torch.cuda.synchronize()
    assert same_bits(buf["score"].cpu().numpy(), g["cem_score"])
    assert np.array_equal(buf["target"].cpu().numpy(), g["cem_target"])


@pytest.mark.parametrize("name", SCENES)
def test_fused_sam_cem_equals_separate_launches(H, torch, ctx, name):
    g = golden(name)
    cube, b, h, w = load_scene(H, torch, g)
    ctx.set_library(g["library"], g["norms"])
    upload_golden_stats(ctx, g)
    p = h * w
    a = device_buffers(torch, p)
    c = device_buffers(torch, p)
    ctx.score_sam_cem(cube, h, w, b, a["score"], a["target"], c["score"], c["target"])
    s = device_buffers(torch, p)
    e = device_buffers(torch, p)
    ctx.score(H.Detector.SAM, cube, h, w, b, s["score"], s["target"])
    ctx.score(H.Detector.CEM, cube, h, w, b, e["score"], e["target"])
    torch.cuda.synchronize()
    for x, y in ((a, s), (c, e)):
        assert same_bits(x["score"].cpu().numpy(), y["score"].cpu().numpy())
        assert np.array_equal(x["target"].cpu().numpy(), y["target"].cpu().numpy())


@pytest.mark.parametrize("name", SCENES)
def test_rx_ace_with_reference_statistics(H, torch, ctx, name):
    g = golden(name)
    cube, b, h, w = load_scene(H, torch, g)
    ctx.set_library(g["library"], g["norms"])
    upload_golden_stats(ctx, g)
    p = h * w
    rx = device_buffers(torch, p)
    ctx.score(H.Detector.RX, cube, h, w, b, rx["score"])
    ace = device_buffers(torch, p)
    ctx.score(H.Detector.ACE, cube, h, w, b, ace["score"], ace["target"], ace["aux"])
    torch.cuda.synchronize()
    e_rx = rel_err(rx["score"].cpu().numpy(), g["rx_score"], FLOOR["rx"])
    e_aux = rel_err(ace["aux"].cpu().numpy(), g["ace_rx"], FLOOR["rx"])
    e_ace = rel_err(ace["score"].cpu().numpy(), g["ace_score"], FLOOR["ace"])
    print(f"{name}: RX max rel {e_rx.max():.3g}, ACE max {e_ace.max():.3g}")
    assert e_rx.max() <= REL and e_aux.max() <= REL and e_ace.max() <= REL
    got_t = ace["target"].cpu().numpy()
    bad = np.nonzero(got_t != g["ace_target"])[0]
    assert bad.size <= max(1, p // 10000), f"{bad.size} ACE target mismatches"


@pytest.mark.parametrize("name", SCENES)
def test_rx_ace_fp32_form_with_reference_statistics(H, torch, ctx, name):
    """The FP32 form (whitening and ACE products on FP32 matrix cores, FP32
    accumulation) against the reference's double-precision scores, on the
    reference's statistics: within the plain bar of 1e-5 relative (floor 1),
    the ACE target as the plain form's (at most one pixel in 10^4 differs)."""
    g = golden(name)
    cube, b, h, w = load_scene(H, torch, g)
    ctx.set_library(g["library"], g["norms"])
    upload_golden_stats(ctx, g)
    p = h * w
    rx = device_buffers(torch, p)
    ctx.score(H.Detector.RX, cube, h, w, b, rx["score"], precision=H.FP32)
    ace = device_buffers(torch, p)
    ctx.score(H.Detector.ACE, cube, h, w, b, ace["score"], ace["target"], ace["aux"], precision=H.FP32)
    torch.cuda.synchronize()
    e_rx = rel_err(rx["score"].cpu().numpy(), g["rx_score"], FLOOR["rx"])
    e_aux = rel_err(ace["aux"].cpu().numpy(), g["ace_rx"], FLOOR["rx"])
    e_ace = rel_err(ace["score"].cpu().numpy(), g["ace_score"], FLOOR["ace"])
    print(f"{name}: FP32 form RX max rel {e_rx.max():.3g}, ACE max {e_ace.max():.3g}")
    assert e_rx.max() <= REL and e_aux.max() <= REL and e_ace.max() <= REL
    bad = np.nonzero(ace["target"].cpu().numpy() != g["ace_target"])[0]
    assert bad.size <= max(1, p // 10000), f"{bad.size} ACE target mismatches"
    # SAM and CEM have no FP32 form
    with pytest.raises(H.HsiError):
        ctx.score(H.Detector.SAM, cube, h, w, b, rx["score"], rx["target"], precision=H.FP32)


@pytest.mark.parametrize("name", SCENES)
def test_moments_and_statistics_built_on_device(H, torch, ctx, oracle_lib, name):
    g = golden(name)
    cube, b, h, w = load_scene(H, torch, g)
    t = g["library"].shape[0]
    ctx.set_library(g["library"], g["norms"])
    ctx.stats_accumulate(cube, h, w, b)
    ctx.stats_build(float(g["loading"]))
    ctx.stats_check()
    st = ctx.stats_read(b, t)
    assert st["count"][0] == g["m_count"]
    lower = np.tril_indices(b)
    assert rel_err(st["sum"], g["m_sum"], 1e-30).max() <= 1e-12
    assert rel_err(st["sum_outer"][lower], g["m_outer"][lower], 1e-30).max() <= 1e-12
    # Statistics from the device moments vs the reference's: float outputs
    # agree to a few ulps (the double internals differ only by sum order).
    for key, want, floor in (("mean", g["st_mean"], 1e-30),
                             ("whitener", g["st_whitener"], None),
                             ("target_white", g["st_target_white"], None),
                             ("target_norm2", g["st_target_norm2"], 1e-30),
                             ("cem_weight", g["st_cem_weight"], None)):
        fl = floor if floor is not None else float(np.abs(want).max()) * 1e-3
        e = rel_err(st[key], want, fl).max()
        print(f"{name}: {key} max rel err {e:.3g}")
        assert e <= 1e-5, key


@pytest.mark.parametrize("name", SCENES)
def test_end_to_end_scores_with_device_statistics(H, torch, ctx, name):
    g = golden(name)
    cube, b, h, w = load_scene(H, torch, g)
    p = h * w
    ctx.set_library(g["library"], g["norms"])
    ctx.stats_accumulate(cube, h, w, b)
    ctx.stats_build(float(g["loading"]))
    ctx.stats_check()
    cem = device_buffers(torch, p)
    ctx.score(H.Detector.CEM, cube, h, w, b, cem["score"], cem["target"])
    rx = device_buffers(torch, p)
    ctx.score(H.Detector.RX, cube, h, w, b, rx["score"])
    ace = device_buffers(torch, p)
    ctx.score(H.Detector.ACE, cube, h, w, b, ace["score"], ace["target"], ace["aux"])
    torch.cuda.synchronize()
    for key, got in (("cem", cem["score"]), ("rx", rx["score"]), ("ace", ace["score"])):
        e = rel_err(got.cpu().numpy(), g[f"{key}_score"], FLOOR[key])
        print(f"{name}: end-to-end {key} max rel {e.max():.3g}")
        assert e.max() <= REL, key


def test_end_to_end_detection_lists(H, torch, ctx, oracle_lib):
    """Cube -> statistics -> scores -> detection lists on each side (device
    statistics against the oracle's build_statistics) for a 640 x 600 x 113
    generator frame with 8 targets. SAM lists are identical; CEM lists are
    identical as (pixel, target) sets except at pixels whose oracle score is
    within 1e-5 of the threshold or of a neighbour inside the NMS window,
    which are counted separately."""
    h, w, b, t = 640, 600, 113, 8
    cfg = H.synth_config(h, w, b, t, seed=20261203 + 91)
    host = H.synth_frame(cfg, 0, threads=8)
    lib = np.asarray(H.synth_library(cfg), np.float32)
    p = h * w
    cube = torch.from_numpy(host).cuda()
    ctx.set_library(lib)
    ctx.stats_accumulate(cube, h, w, b)
    ctx.stats_build(1e-6)
    ctx.stats_check()
    sam, cem = device_buffers(torch, p), device_buffers(torch, p)
    ctx.score_sam_cem(cube, h, w, b, sam["score"], sam["target"], cem["score"], cem["target"])
    torch.cuda.synchronize()
    count, s1, s2 = oracle_lib.moments(host, pixels=p, threads=8)
    st = oracle_lib.build_statistics(count, s1, s2, lib, 1e-6)
    want = {"sam": oracle_lib.sam_best(host, lib, pixels=p, threads=8),
            "cem": oracle_lib.cem_best(host, st, pixels=p, threads=8)}
    got = {"sam": (sam["score"].cpu().numpy(), sam["target"].cpu().numpy()),
           "cem": (cem["score"].cpu().numpy(), cem["target"].cpu().numpy())}
    cuts = {"sam": (0.10, False), "cem": (0.5, True)}
    for det in ("sam", "cem"):
        thr, higher = cuts[det]
        gl, gfound = run_detect(H, torch, ctx, got[det][0], got[det][1], h, w, thr, 2, higher)
        wl, wfound = oracle_lib.detect(want[det][0], want[det][1], h, w, thr, 2, higher)
        gset = set(zip((gl["y"] * w + gl["x"]).tolist(), gl["target"].tolist()))
        wset = set(zip((wl["y"] * w + wl["x"]).tolist(), wl["target"].tolist()))
        if det == "sam":
            assert gfound == wfound and gset == wset
            continue
        near = near_cut(want[det][0], h, w, thr, 2, REL)
        off = [(q, tt) for q, tt in gset ^ wset if not near[q]]
        print(f"end-to-end {det}: {len(gset)} device / {len(wset)} oracle detections, "
              f"{len(gset ^ wset) - len(off)} differ at near-threshold or near-tie pixels")
        assert not off, off[:10]
        assert len(wset) > 100


def test_singular_covariance_is_reported_with_loading(H, torch, ctx):
    g = golden("edge_cases")
    cube = H.to_device_cube(g["sing_cube"], 10, 10)
    ctx.stats_accumulate(cube, 10, 10, 4)
    ctx.stats_build(0.0)
    with pytest.raises(H.HsiError) as info:
        ctx.stats_check()
    assert info.value.status == H.ERR_NOT_POSITIVE_DEFINITE
    assert "loading" in str(info.value)
    ctx.stats_build(1e-3)
    ctx.stats_check()
    st = ctx.stats_read(4, 0)
    want = g["sing_l3_whitener"]
    assert rel_err(st["whitener"], want, float(np.abs(want).max()) * 1e-3).max() <= 1e-5


def test_near_singular_covariance_is_rejected_with_its_own_status(H, torch, ctx):
    """The reference factorises this covariance (its smallest pivot is ~1e-15
    of its diagonal, positive); the device rejects pivots at or below
    4*256*eps of their diagonal and says so in the status word and message.
    With loading both factorise and agree."""
    g = golden("edge_cases")
    cube = H.to_device_cube(g["nsing_cube"], 16, 16)
    ctx.stats_accumulate(cube, 16, 16, 6)
    ctx.stats_build(0.0)
    cov, corr = ctx.stats_status()
    assert cov == H.STATUS_COV_NOT_PD | H.STATUS_BELOW_TOLERANCE, cov
    with pytest.raises(H.HsiError) as info:
        ctx.stats_check()
    assert info.value.status == H.ERR_NOT_POSITIVE_DEFINITE
    assert "loading" in str(info.value) and "4*256*eps" in str(info.value)
    ctx.stats_build(1e-3)
    ctx.stats_check()
    assert ctx.stats_status() == [0, 0]
    st = ctx.stats_read(6, 0)
    want = g["nsing_l3_whitener"]
    assert rel_err(st["whitener"], want, float(np.abs(want).max()) * 1e-3).max() <= 1e-5
    assert rel_err(st["mean"], g["nsing_l3_mean"], 1.0).max() <= 1e-6


def run_detect(H, torch, ctx, score, target, h, w, thr, nms, higher, cap=None):
    p = h * w
    cap = p if cap is None else cap
    s = torch.from_numpy(np.ascontiguousarray(score, np.float32).reshape(-1)).cuda()
    t = None if target is None else torch.from_numpy(np.ascontiguousarray(target, np.int32).reshape(-1)).cuda()
    out = torch.zeros((max(cap, 1), 4), dtype=torch.int32, device="cuda")
    count = torch.full((1,), 12345, dtype=torch.int32, device="cuda")
    ctx.detect(s, t, h, w, H.DetectionParams(thr, nms, cap, higher), out, count)
    torch.cuda.synchronize()
    found = int(count.item())
    return H.detections_to_numpy(out, min(found, cap)), found


@pytest.mark.parametrize("name", SCENES)
@pytest.mark.parametrize("det", ["sam", "rx", "ace", "cem"])
@pytest.mark.parametrize("nms", [0, 2])
def test_detect_matches_reference_lists(H, torch, ctx, name, det, nms):
    g = golden(name)
    _, h, w = g["cube"].shape
    target = None if det == "rx" else g[f"{det}_target"]
    dets, found = run_detect(H, torch, ctx, g[f"{det}_score"], target, h, w,
                             float(g[f"{det}_threshold"]), nms, det != "sam")
    want = g[f"{det}_det{nms}_xyt"]
    assert found == want.shape[0]
    assert np.array_equal(np.stack([dets["x"], dets["y"], dets["target"]], 1).reshape(-1, 3), want)
    assert same_bits(dets["score"], g[f"{det}_det{nms}_score"])


@pytest.mark.parametrize("case", ["blob", "plateau", "pol_low", "pol_high", "extremum",
                                  "ties_low", "ties_high"])
def test_detect_edge_maps(H, torch, ctx, case):
    g = golden("edge_cases")
    mp = g[f"det_{case}_map"]
    thr, nms, higher = g[f"det_{case}_args"]
    h, w = mp.shape
    tg = (np.arange(h * w) % 3).astype(np.int32)
    dets, found = run_detect(H, torch, ctx, mp, tg, h, w, float(thr), int(nms), bool(higher))
    want = g[f"det_{case}_xyt"]
    assert found == want.shape[0]
    assert np.array_equal(np.stack([dets["x"], dets["y"], dets["target"]], 1).reshape(-1, 3), want)


@pytest.mark.parametrize("nms,higher", [(1, True), (2, False), (3, True), (5, False)])
def test_detect_large_ragged_map(H, torch, ctx, oracle_lib, nms, higher):
    """A 1000 x 1037 map: 4051 workgroups of 256 pixels, the last one
    ragged (200 pixels), with neighbourhood runs that wrap across image rows.
    Scores are quantised to 1/64 so plateaus of equal neighbours exercise the
    raster-order tie rule; about 5% of pixels pass."""
    rng = np.random.default_rng(11 + nms)
    h, w = 1000, 1037
    score = (np.round(rng.random((h, w)) * 64) / 64).astype(np.float32)
    target = rng.integers(0, 8, size=h * w).astype(np.int32)
    thr = 0.97 if higher else 0.03
    want, found_cpu = oracle_lib.detect(score, target, h, w, thr, nms, higher)
    dets, found = run_detect(H, torch, ctx, score, target, h, w, thr, nms, higher)
    assert found == found_cpu > 1000
    for k in ("x", "y", "target"):
        assert np.array_equal(dets[k], want[k])
    assert same_bits(dets["score"], want["score"])


@pytest.mark.parametrize("h,w,nms,higher", [(2048, 1024, 1, True), (2049, 1024, 2, False), (2111, 997, 3, True)])
def test_detect_group_rank_boundaries(H, torch, ctx, oracle_lib, h, w, nms, higher):
    """For radii 1-3 the flag pass counts per 1024-pixel group and, up to
    2048 groups, each write workgroup finds its first slot from the counts
    of the groups before it; past that a scan over the group counts runs
    first. 2048 x 1024 is exactly 2048 groups, 2049 x 1024 is 2049 (the
    scan), and 2111 x 997 is 2056 with a ragged last group of 347 pixels and
    neighbourhood runs that wrap across rows. Lists equal detect_cpu's
    element by element, also when the capacity cuts them short."""
    rng = np.random.default_rng(101 + nms)
    score = (np.round(rng.random((h, w)) * 64) / 64).astype(np.float32)
    target = rng.integers(0, 8, size=h * w).astype(np.int32)
    thr = 0.97 if higher else 0.03
    want, found_cpu = oracle_lib.detect(score, target, h, w, thr, nms, higher)
    dets, found = run_detect(H, torch, ctx, score, target, h, w, thr, nms, higher)
    assert found == found_cpu > 10000
    for k in ("x", "y", "target"):
        assert np.array_equal(dets[k], want[k])
    assert same_bits(dets["score"], want["score"])
    cap = found_cpu // 3
    dets, found = run_detect(H, torch, ctx, score, target, h, w, thr, nms, higher, cap=cap)
    assert found == found_cpu and len(dets) == cap
    for k in ("x", "y", "target"):
        assert np.array_equal(dets[k], want[k][:cap])


def test_detect_saturates_without_overrun(H, torch, ctx, oracle_lib):
    rng = np.random.default_rng(3)
    h, w = 300, 500
    score = rng.random((h, w), dtype=np.float32)
    target = rng.integers(0, 4, size=h * w).astype(np.int32)
    want, found_cpu = oracle_lib.detect(score, target, h, w, 0.5, 0, False)
    cap = 1000
    dets, found = run_detect(H, torch, ctx, score, target, h, w, 0.5, 0, False, cap=cap)
    assert found == found_cpu > cap
    assert len(dets) == cap
    assert np.array_equal(dets["x"], want["x"][:cap]) and np.array_equal(dets["y"], want["y"][:cap])
    # an empty result: nothing passes
    dets, found = run_detect(H, torch, ctx, score, target, h, w, -1.0, 2, False, cap=16)
    assert found == 0


# ---- precision form ---------------------------------------------------------------
MIXED_REL = 1e-3  # score error bar of the mixed form against the FP64 oracle


@pytest.mark.parametrize("name", SCENES)
@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_mixed_rx_ace_against_fp64_oracle(H, torch, ctx, oracle_lib, name, dtype):
    g = golden(name)
    b, h, w = g["cube"].shape
    p = h * w
    src = g["cube"] if dtype == "f32" else g["cube"].astype(np.float16).astype(np.float32)
    cube = H.to_device_cube(src, h, w, dtype=dtype)
    ctx.set_library(g["library"], g["norms"])
    upload_golden_stats(ctx, g)
    st = oracle_lib.Statistics(b, g["library"].shape[0])
    st.mean, st.whitener = g["st_mean"], g["st_whitener"]
    st.target_white, st.target_norm2 = g["st_target_white"], g["st_target_norm2"]
    st.cem_weight = g["st_cem_weight"]
    want_ace, want_t, want_rx = oracle_lib.ace_best(src.reshape(b, -1), st)
    rx = device_buffers(torch, p)
    ctx.score(H.Detector.RX, cube, h, w, b, rx["score"], precision=H.MIXED)
    ace = device_buffers(torch, p)
    ctx.score(H.Detector.ACE, cube, h, w, b, ace["score"], ace["target"], ace["aux"],
              precision=H.MIXED)
    torch.cuda.synchronize()
    e_rx = rel_err(rx["score"].cpu().numpy(), want_rx, FLOOR["rx"]).max()
    e_aux = rel_err(ace["aux"].cpu().numpy(), want_rx, FLOOR["rx"]).max()
    e_ace = rel_err(ace["score"].cpu().numpy(), want_ace, FLOOR["ace"]).max()
    print(f"{name} {dtype}: mixed RX max rel {e_rx:.3g}, ACE max {e_ace:.3g}")
    assert e_rx <= MIXED_REL and e_aux <= MIXED_REL and e_ace <= MIXED_REL
    agree = (ace["target"].cpu().numpy() == want_t).mean()
    assert agree >= 0.99


@pytest.mark.parametrize("name", SCENES)
def test_mixed_moments_against_oracle(H, torch, ctx, name):
    """FP16 hi/lo products, FP32 tiles, FP64 across tiles: the moments and the
    scores built on them stay within the precision form's bar."""
    g = golden(name)
    cube, b, h, w = load_scene(H, torch, g)
    p = h * w
    ctx.set_library(g["library"], g["norms"])
    ctx.stats_accumulate(cube, h, w, b, precision=H.MIXED)
    ctx.stats_build(float(g["loading"]))
    ctx.stats_check()
    st = ctx.stats_read(b, g["library"].shape[0])
    assert st["count"][0] == g["m_count"]
    lower = np.tril_indices(b)
    e_sum = rel_err(st["sum"], g["m_sum"], 1e-30).max()
    e_outer = rel_err(st["sum_outer"][lower], g["m_outer"][lower], 1e-30).max()
    scores = {}
    for det in (H.Detector.CEM, H.Detector.RX, H.Detector.ACE):
        buf = device_buffers(torch, p)
        ctx.score(det, cube, h, w, b, buf["score"], buf["target"], buf["aux"])
        scores[det] = buf["score"]
    torch.cuda.synchronize()
    errs = {k: rel_err(scores[d].cpu().numpy(), g[f"{k}_score"], FLOOR[k]).max()
            for k, d in (("cem", H.Detector.CEM), ("rx", H.Detector.RX), ("ace", H.Detector.ACE))}
    print(f"{name}: mixed moments sum {e_sum:.3g} outer {e_outer:.3g}; scores {errs}")
    assert e_sum <= 1e-6 and e_outer <= 1e-6
    assert max(errs.values()) <= MIXED_REL
