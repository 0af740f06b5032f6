"""Batched uneven scenes (one persistent launch per shape class) against the
single-scene path, scene by scene: identical bits, before and after the
autotuner picks launch shapes; and the device demosaic against the
snapshot-mosaic layout it inverts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import torch
    import hsi_mi355x
    assert torch.cuda.is_available()
    return hsi_mi355x


@pytest.mark.parametrize("raw_type", ["u8", "u16", "f32"])
def test_demosaic_matches_mosaic_layout(H, raw_type):
    import torch
    rng = np.random.default_rng(3)
    period, rh, rw, bands = 4, 36, 52, 16
    if raw_type == "u8":
        raw = rng.integers(0, 256, (rh, rw), dtype=np.uint8)
        scale, code = np.float32(1.0) / np.float32(255.0), H.RAW_U8
    elif raw_type == "u16":
        raw = rng.integers(0, 65536, (rh, rw), dtype=np.uint16)
        scale, code = np.float32(1.0) / np.float32(65535.0), H.RAW_U16
    else:
        raw = rng.random((rh, rw), dtype=np.float32)
        scale, code = np.float32(1.0), H.RAW_F32
    order = rng.permutation(period * period).astype(np.int32)[:bands]
    h, w = rh // period, rw // period
    stride = H.plane_stride(h, w)
    cube = torch.full((bands, stride), -1.0, dtype=torch.float32, device="cuda")
    raw_d = torch.from_numpy(raw.astype(raw.dtype) if raw_type != "u16" else raw.view(np.int16)).cuda()
    ctx = H.Context(0)
    print(f"demosaic {raw_type} {rh}x{rw} period {period}", flush=True)
    H.demosaic(ctx, raw_d, code, rh, rw, period, bands, cube, band_order=order)
    torch.cuda.synchronize()
    got = cube.cpu().numpy()[:, :h * w].reshape(bands, h, w)
    for b in range(bands):
        oy, ox = divmod(int(order[b]), period)
        want = raw[oy::period, ox::period][:h, :w].astype(np.float32) * scale
        assert np.array_equal(got[b], want), b
    ctx.close()


def make_scenes(H, torch, n, seed):
    rng = np.random.default_rng(seed)
    scenes, keep = [], []
    for i in range(n):
        b = int(rng.choice([16, 113, 224]))
        h, w = int(rng.integers(9, 70)), int(rng.integers(7, 90))
        t = int(rng.integers(1, 33))
        cube = rng.uniform(0.02, 0.9, (b, h * w)).astype(np.float32)
        lib = rng.uniform(0.05, 0.9, (t, b)).astype(np.float32)
        # plant a few library pixels so detections are non-trivial
        for k in range(min(t, 3)):
            cube[:, (k * 7) % (h * w)] = lib[k] * 0.8
        dev = H.to_device_cube(cube.reshape(b, h, w), h, w)
        p = h * w
        f = dict(dtype=torch.float32, device="cuda")
        it = dict(dtype=torch.int32, device="cuda")
        outs = {k: torch.full((p,), float("nan"), **f) for k in ("sam_score", "cem_score", "rx", "ace_score")}
        outs.update({k: torch.full((p,), -7, **it) for k in ("sam_target", "cem_target", "ace_target")})
        scenes.append(dict(cube=dev, height=h, width=w, library=lib, **outs))
    return scenes


def single(H, torch, sc):
    ctx = H.Context(0)
    ctx.set_library(sc["library"])
    h, w, b = sc["height"], sc["width"], sc["cube"].shape[0]
    p = h * w
    ctx.stats_accumulate(sc["cube"], h, w, b)
    ctx.stats_build(1e-6)
    f = dict(dtype=torch.float32, device="cuda")
    it = dict(dtype=torch.int32, device="cuda")
    r = {k: torch.empty(p, **f) for k in ("sam_score", "cem_score", "rx", "ace_score")}
    r.update({k: torch.empty(p, **it) for k in ("sam_target", "cem_target", "ace_target")})
    ctx.score_sam_cem(sc["cube"], h, w, b, r["sam_score"], r["sam_target"], r["cem_score"], r["cem_target"])
    ctx.score(H.Detector.ACE, sc["cube"], h, w, b, r["ace_score"], r["ace_target"], r["rx"])
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in r.items()}
    ctx.close()
    return out


def test_batch_equals_single_scene_path(H):
    import torch
    scenes = make_scenes(H, torch, 12, seed=21)
    batch = H.Batch(0)
    batch.set_scenes(scenes)
    batch.statistics(1e-6)
    batch.check()
    print("batch of 12 scenes", [(s["cube"].shape[0], s["height"], s["width"], s["library"].shape[0])
                                 for s in scenes], flush=True)
    batch.score()
    torch.cuda.synchronize()
    keys = ("sam_score", "sam_target", "cem_score", "cem_target", "rx", "ace_score", "ace_target")
    first = {i: {k: s[k].cpu().numpy().copy() for k in keys} for i, s in enumerate(scenes)}
    for i, sc in enumerate(scenes):
        want = single(H, torch, sc)
        for k in keys:
            assert np.array_equal(first[i][k].view(np.uint32), want[k].view(np.uint32)), (i, k)
    report = batch.autotune()
    print(report)
    assert "sam/cem" in report and "rx/ace" in report
    for s in scenes:
        for k in keys:
            s[k].fill_(-3)
    batch.score()
    torch.cuda.synchronize()
    for i, s in enumerate(scenes):
        for k in keys:
            assert np.array_equal(s[k].cpu().numpy().view(np.uint32), first[i][k].view(np.uint32)), (i, k)
    batch.close()


KEYS = ("sam_score", "sam_target", "cem_score", "cem_target", "rx", "ace_score", "ace_target")


@pytest.fixture(scope="module")
def uneven_batch(H):
    """The 64-scene uneven workload (hsi_mi355x.workloads, the bench's seed):
    the 16 snapshot mosaics demosaiced on the device, statistics and scores
    from one eager batch run."""
    import os
    import torch
    from hsi_mi355x.workloads import make_scene, uneven_scene_specs
    ctx0 = H.Context(0)
    specs = uneven_scene_specs(20261203 + 4000)
    threads = min(16, os.cpu_count() or 1)
    scenes = [make_scene(H, torch, ctx0, sp, threads=threads) for sp in specs]
    print(f"64 scenes, {sum(s['height'] * s['width'] for s in scenes) / 1e6:.1f} Mpix", flush=True)
    batch = H.Batch(0)
    batch.set_scenes(scenes)
    batch.run(1e-6, use_graph=False)
    batch.check()
    torch.cuda.synchronize()
    eager = [{k: s[k].cpu().numpy().copy() for k in KEYS} for s in scenes]
    yield dict(batch=batch, scenes=scenes, specs=specs, eager=eager, ctx0=ctx0)
    batch.close()
    ctx0.close()


def test_uneven_batch_against_the_oracle_on_stripes(H, uneven_batch, oracle_lib):
    """Every scene's SAM, CEM, RX and ACE on its first 64 rows against the
    oracle, which scores the same device cube (demosaiced ones included) with
    the batch's statistics: angles within 1e-6 rad, scores within 1e-5
    relative (floor 1), targets equal."""
    import os
    from gpu_util import ANGLE_TOL, REL, rel_err
    threads = min(16, os.cpu_count() or 1)
    batch = uneven_batch["batch"]
    for i, (sc, out) in enumerate(zip(uneven_batch["scenes"], uneven_batch["eager"])):
        h, w, b = sc["height"], sc["width"], sc["cube"].shape[0]
        rows = min(64, h)
        p = rows * w
        cube = np.ascontiguousarray(sc["cube"][:, :p].cpu().numpy()).reshape(b, rows, w)
        lib = sc["library"]
        t = lib.shape[0]
        dev = batch.context(i).stats_read(b, t)
        st = oracle_lib.Statistics(b, t)
        st.mean[:] = dev["mean"]
        st.whitener[:] = dev["whitener"]
        st.target_white[:] = dev["target_white"][:t]
        st.target_norm2[:] = dev["target_norm2"][:t]
        st.cem_weight[:] = dev["cem_weight"][:t]
        st.status = oracle_lib.STATS_OK
        ws, wt = oracle_lib.sam_best(cube, lib, threads=threads)
        assert np.abs(out["sam_score"][:p].astype(np.float64) - ws).max() <= ANGLE_TOL, i
        assert np.array_equal(out["sam_target"][:p], wt), i
        cs, ct = oracle_lib.cem_best(cube, st, threads=threads)
        assert rel_err(out["cem_score"][:p], cs, 1.0).max() <= REL, i
        assert (out["cem_target"][:p] == ct).mean() >= 0.9999, i
        acs, act, arx = oracle_lib.ace_best(cube, st, threads=threads)
        assert rel_err(out["rx"][:p], arx, 1.0).max() <= REL, i
        assert rel_err(out["ace_score"][:p], acs, 1.0).max() <= REL, i
        assert (out["ace_target"][:p] == act).mean() >= 0.9999, i
    print("64 scenes: stripes match the oracle")


def test_uneven_batch_equals_single_scene_path_at_full_size(H, uneven_batch):
    import torch
    for i, (sc, out) in enumerate(zip(uneven_batch["scenes"], uneven_batch["eager"])):
        want = single(H, torch, sc)
        for k in KEYS:
            assert np.array_equal(out[k].view(np.uint32), want[k].view(np.uint32)), (i, k)


def test_uneven_batch_graph_replays_equal_the_eager_run(H, uneven_batch):
    """The captured statistics + scoring graph, replayed three times with the
    program's own eager work between replays, reproduces the eager run bit
    for bit each time."""
    import torch
    batch, scenes, eager = uneven_batch["batch"], uneven_batch["scenes"], uneven_batch["eager"]
    for replay in range(3):
        for s in scenes:
            for k in KEYS:
                s[k].fill_(-3)
        junk = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work
        batch.run(1e-6, use_graph=True)
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for i, s in enumerate(scenes):
            for k in KEYS:
                assert np.array_equal(s[k].cpu().numpy().view(np.uint32), eager[i][k].view(np.uint32)), (replay, i, k)
    batch.check()
