#!/usr/bin/env python3
"""Frame-scoring benchmark on one MI355X.

Workload (the plain-FP32 video configuration): frames of 1024x1024 pixels x
113 bands, BSQ fp32, drawn by the seeded generator (Dirichlet mix of 8
endmembers, per-frame drift of the mix, 8-target library planted at 0.5%
each, gain U(0.3,1.2), noise N(0,0.005)). One step = one frame through the
whole per-frame path, with the frame already resident in HBM:
    moments (FP64 matrix cores) -> statistics build (FP64)
    -> fused SAM+CEM (one read of the cube) -> two threshold+NMS+compact
Three frames are in flight (a context and a stream per slot, as the
reference Pipeline's default of three), so one frame's HBM-bound SAM+CEM and
single-CU build overlap the next frame's moments; --in-flight 1 gives the
serial form and --separate-sam scores SAM on a second stream before the
statistics. `value` is frames x pixels / wall time of the timed steps
(Mpix/s).

Per-kernel rates come from HIP events on the launch stream, averaged over
repeats, and the roofline entry is for the kernel that takes the largest
share of a step. The CPU baseline times the reference's own CPU scorer
(oracle/_ref, built from the reference sources) - or the oracle restatement
when that is absent - on a bounded sample of the same frame.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "hsi-detectors-cuda_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP16_MFMA_PEAK_TFLOPS = 2500.0

H_, W_, B_, T_ = 1024, 1024, 113, 8

# HBM bytes per launch of each candidate dominant kernel on this workload,
# from rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by a 1 GiB copy (FETCH
# KiB x 2.0, WRITE KiB x 1.0); the runs and the table are recorded in
# profiles/hbm_traffic_1024x1024x113.md.
MEASURED_TRAFFIC_BYTES = {
    "moments_f64_kernel": 231619.9 * 2.0 * 1024 + 29600.0 * 1024,
    "sam_cem_fused_kernel": 232832.0 * 2.0 * 1024 + 18432.0 * 1024,
}
LOADING = 1e-6
SAM_THRESHOLD = 0.10
CEM_THRESHOLD = 0.5


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--frames", type=int, default=8, help="distinct resident frames cycled")
    ap.add_argument("--in-flight", type=int, default=3,
                    help="frames in flight (slots of two streams and two contexts each); 3 is the reference Pipeline default")
    ap.add_argument("--repeats", type=int, default=20, help="per-kernel timing repeats")
    ap.add_argument("--separate-sam", action="store_true",
                    help="score SAM on its own stream (before the statistics) instead of fused with CEM")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rows", type=int, default=1024, help="rows of a frame in the CPU sample")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 64-scene batched leg")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned pre-screen leg (2048x2048 frames)")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 224-band precision-form leg (plain vs mixed RX/ACE)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB buffer 5 times (reference kernel for HBM byte counters)")
    return ap.parse_args()


def copy_calibrate(mib):
    """A device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    readings give the factor that turns those counters into bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for i in range(5):
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy {i}: {mib} MiB read, {mib} MiB written")
    print(json.dumps({"copy_mib": mib, "copies": 5}), flush=True)


def time_on_stream(torch, fn, repeats):
    """Average ms of fn() over `repeats`, bracketed by events on the current stream."""
    fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(repeats):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / repeats


def cpu_baseline(host_frame, library, rows):
    """Reference CPU scorer (or its restatement) on `rows` rows of a frame."""
    import ctypes as C
    sample_px = rows * W_
    sub = np.ascontiguousarray(host_frame[:, :sample_px])
    ref_path = os.path.join(REPO, "oracle", "_ref", "libhsiref.so")
    if os.path.exists(ref_path):
        kind = "reference"
        ref = C.CDLL(ref_path)
        f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        ref.ref_library_norms.argtypes = [f32, C.c_int, C.c_int, f32]
        ref.ref_sam_best.argtypes = [f32, C.c_int, C.c_int, C.c_int, f32, f32, C.c_int, f32, i32]
        ref.ref_build_statistics.restype = C.c_int
        ref.ref_build_statistics.argtypes = [f32, C.c_int, C.c_int, C.c_int, f32, C.c_int, C.c_int,
                                             C.c_double, f32, f32, f32, f32, f32, C.c_char_p, C.c_int]
        ref.ref_cem_best.argtypes = [f32, C.c_int, C.c_int, C.c_int, f32, f32, f32, f32, f32,
                                     C.c_int, f32, i32]
        ref.ref_detect.restype = C.c_int64
        ref.ref_detect.argtypes = [f32, i32, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, i32,
                                   f32, C.c_int64]
        cap = 4096
        det_xyt = np.empty(3 * cap, np.int32)
        det_score = np.empty(cap, np.float32)
        norms = np.empty(T_, np.float32)
        ref.ref_library_norms(library, T_, B_, norms)
        mean = np.empty(B_, np.float32)
        white = np.empty(B_ * B_, np.float32)
        tw = np.empty(T_ * B_, np.float32)
        tn = np.empty(T_, np.float32)
        cw = np.empty(T_ * B_, np.float32)
        err = C.create_string_buffer(256)
        score = np.empty(sample_px, np.float32)
        target = np.empty(sample_px, np.int32)
        t0 = time.perf_counter()
        ok = ref.ref_build_statistics(sub, rows, W_, B_, library, T_, B_, LOADING, mean, white, tw,
                                      tn, cw, err, 256)
        t1 = time.perf_counter()
        ref.ref_sam_best(sub, rows, W_, B_, library, norms, T_, score, target)
        ref.ref_detect(score, target, rows, W_, SAM_THRESHOLD, 2, 0, det_xyt, det_score, cap)
        t2 = time.perf_counter()
        ref.ref_cem_best(sub, rows, W_, B_, mean, white, tw, tn, cw, T_, score, target)
        ref.ref_detect(score, target, rows, W_, CEM_THRESHOLD, 2, 1, det_xyt, det_score, cap)
        t3 = time.perf_counter()
        if not ok:
            raise RuntimeError(err.value.decode())
    else:
        kind = "port"
        import oracle
        t0 = time.perf_counter()
        count, s, so = oracle.moments(sub)
        st = oracle.build_statistics(count, s, so, library, LOADING)
        t1 = time.perf_counter()
        score, target = oracle.sam_best(sub, library)
        oracle.detect(score, target, rows, W_, SAM_THRESHOLD, 2, False, capacity=4096)
        t2 = time.perf_counter()
        score, target = oracle.cem_best(sub, st)
        oracle.detect(score, target, rows, W_, CEM_THRESHOLD, 2, True, capacity=4096)
        t3 = time.perf_counter()
    total = t3 - t0
    return {
        "value": round(sample_px / total / 1e6, 4),
        "unit": "Mpix/s",
        "cores": 1,
        "kind": kind,
        "sample": f"rows 0-{rows - 1} of frame 0 ({rows}x{W_}x{B_}, {T_} targets): moments+"
                  f"build_statistics, sam_best_cpu+detect_cpu, cem_best_cpu+detect_cpu, "
                  f"single thread (the reference scorers are serial)",
        "seconds": round(total, 3),
        "per_detector_mpix_s": {"stats": round(sample_px / (t1 - t0) / 1e6, 4),
                                "sam": round(sample_px / (t2 - t1) / 1e6, 4),
                                "cem": round(sample_px / (t3 - t2) / 1e6, 4)},
    }


def precision_form(H, torch, repeats):
    """RX+ACE in the precision form: 1024x1024 px x 224 bands, 16 targets, the
    FP32 frame stored on the device as FP16. Per frame the RX/ACE path is
    moments -> FP64 build -> ACE (RX as its side output).
      plain: FP64 moments, FP64-accumulated whitening on the FP32 cube;
      mixed: FP16 hi/lo moments with FP32 tiles reduced in FP64, FP64 build,
             FP16 MFMA whitening of the FP16 cube (hi/lo whitener), FP64 ratios.
    Error is reported against the plain path (FP64 throughout; the parity
    tests hold it within ~1e-7 of the FP64 oracle), recall as the share of the
    plain path's detections (RX at a chi-squared false-alarm rate of 1e-3, ACE
    at 0.5, NMS radius 2) that the mixed path also reports. The tests check
    the mixed path against the FP64 oracle itself."""
    h, w, b, t = 1024, 1024, 224, 16
    p = h * w
    cfg = H.synth_config(h, w, b, t, seed=20261203 + 3000)
    lib = H.synth_library(cfg)
    host = H.synth_frame(cfg, 0, threads=16)
    c32 = torch.from_numpy(host).cuda()
    c16 = c32.half()
    del host
    ctx = H.Context(0)
    ctx.set_library(lib)
    ctx.reserve(p, b)
    f32 = dict(dtype=torch.float32, device="cuda")
    i32 = dict(dtype=torch.int32, device="cuda")
    ace_s, rx_s, ace_t = torch.empty(p, **f32), torch.empty(p, **f32), torch.empty(p, **i32)

    def plain():
        ctx.stats_accumulate(c32, h, w, b)
        ctx.stats_build(LOADING)
        ctx.score(H.Detector.ACE, c32, h, w, b, ace_s, ace_t, rx_s)

    def mixed():
        ctx.stats_accumulate(c32, h, w, b, precision=H.MIXED)
        ctx.stats_build(LOADING)
        ctx.score(H.Detector.ACE, c16, h, w, b, ace_s, ace_t, rx_s, precision=H.MIXED)

    def mixed32():
        ctx.stats_accumulate(c32, h, w, b, precision=H.MIXED)
        ctx.stats_build(LOADING)
        ctx.score(H.Detector.ACE, c32, h, w, b, ace_s, ace_t, rx_s, precision=H.MIXED)

    def score_only(prec, cube):
        return lambda: ctx.score(H.Detector.ACE, cube, h, w, b, ace_s, ace_t, rx_s, precision=prec)

    cap = 1 << 16
    dets = torch.empty((cap, 4), **i32)
    cnt = torch.empty(1, **i32)

    def detections(score, target, thr):
        ctx.detect(score, target, h, w, H.DetectionParams(thr, 2, cap, True), dets, cnt)
        torch.cuda.synchronize()
        n = min(int(cnt.item()), cap)
        xy = dets[:n, :2].cpu().numpy()
        return set(map(tuple, xy.tolist()))

    rx_thr = H.rx_threshold(b, 1e-3)

    def outputs(fn):
        fn()
        ctx.stats_check()
        torch.cuda.synchronize()
        return (ace_s.cpu().numpy(), ace_t.cpu().numpy(), rx_s.cpu().numpy(),
                detections(rx_s, None, rx_thr), detections(ace_s, ace_t, 0.5))

    ms = {"plain_path": time_on_stream(torch, plain, repeats),
          "mixed_path": time_on_stream(torch, mixed, repeats),
          "plain_score": time_on_stream(torch, score_only(H.PLAIN, c32), repeats)}
    pa, pt, pr, prx, pace = outputs(plain)
    ms["mixed_score"] = time_on_stream(torch, score_only(H.MIXED, c16), repeats)
    ma, mt, mr, mrx, mace = outputs(mixed)
    ms["mixed32_path"] = time_on_stream(torch, mixed32, repeats)
    ms["mixed32_score"] = time_on_stream(torch, score_only(H.MIXED, c32), repeats)
    qa, qt, qr, qrx, qace = outputs(mixed32)
    log("  precision form: " + ", ".join(f"{k} {v * 1e3:.1f} us" for k, v in ms.items()))
    recall = {
        "rx": {"plain_detections": len(prx), "mixed_detections": len(mrx),
               "recall": round(len(prx & mrx) / max(len(prx), 1), 6)},
        "ace": {"plain_detections": len(pace), "mixed_detections": len(mace),
                "recall": round(len(pace & mace) / max(len(pace), 1), 6)},
    }
    rel_rx = np.abs(mr.astype(np.float64) - pr) / np.maximum(np.abs(pr.astype(np.float64)), 1.0)
    rel_rx32 = np.abs(qr.astype(np.float64) - pr) / np.maximum(np.abs(pr.astype(np.float64)), 1.0)
    ctx.close()
    pix = p / 1e6
    return {
        "workload": "RX+ACE, 1024x1024x224, 16 targets, FP32 frame stored as FP16 on the device; "
                    "path = moments + FP64 build + ACE with RX side output",
        "plain": {"path_us": round(ms["plain_path"] * 1e3, 1), "score_us": round(ms["plain_score"] * 1e3, 1),
                  "path_mpix_s": round(pix / (ms["plain_path"] / 1e3), 1),
                  "note": "FP32 cube, FP64 moments and FP64-accumulated whitening"},
        "mixed": {"path_us": round(ms["mixed_path"] * 1e3, 1), "score_us": round(ms["mixed_score"] * 1e3, 1),
                  "path_mpix_s": round(pix / (ms["mixed_path"] / 1e3), 1)},
        "speedup_path": round(ms["plain_path"] / ms["mixed_path"], 3),
        "speedup_score": round(ms["plain_score"] / ms["mixed_score"], 3),
        "recall_vs_plain": recall,
        "error_vs_plain_fp64": {"rx_max_rel": float(f"{rel_rx.max():.3g}"),
                                "ace_max_abs": float(f"{np.abs(ma.astype(np.float64) - pa).max():.3g}"),
                                "ace_target_agree": round(float((mt == pt).mean()), 6),
                                "note": "mostly the FP16 storage of the cube itself"},
        "mixed_fp32_cube": {
            "path_us": round(ms["mixed32_path"] * 1e3, 1), "score_us": round(ms["mixed32_score"] * 1e3, 1),
            "speedup_path": round(ms["plain_path"] / ms["mixed32_path"], 3),
            "speedup_score": round(ms["plain_score"] / ms["mixed32_score"], 3),
            "recall_vs_plain": {"rx": round(len(prx & qrx) / max(len(prx), 1), 6),
                                "ace": round(len(pace & qace) / max(len(pace), 1), 6)},
            "error_vs_plain_fp64": {"rx_max_rel": float(f"{rel_rx32.max():.3g}"),
                                    "ace_max_abs": float(f"{np.abs(qa.astype(np.float64) - pa).max():.3g}")},
            "note": "same arithmetic on the FP32 cube (hi/lo split of the pixels too): no input quantisation"},
    }


def learned_form(H, torch, repeats, size=2048, train_frames=3, eval_frames=3):
    """RX+ACE with the learned pre-screen against the plain full-frame path on
    {size}x{size} x 113-band frames, 8 targets, 1% planted anomalies (a 9th
    endmember at abundance U(0.3, 1.0)). The network is trained on the GPU by
    the library's own kernels from plain labels of `train_frames` frames and
    judged on `eval_frames` held-out frames: recall of the plain detection
    lists (RX at a chi-squared false-alarm rate of 1e-6, ACE at 0.5, NMS
    radius 2) and the false-skip count. Times are HIP-event averages on one
    held-out frame with statistics built: plain = full-frame ACE with the RX
    side output; learned = inference + compaction at the cut + exact
    rescoring of the candidates."""
    from hsi_mi355x.prescreen import Prescreen
    h = w = size
    b, t = 113, 8
    cfg = H.synth_config(h, w, b, t, seed=20261203 + 5000, anomaly_fraction=0.01, drift=0.2,
                         drift_period=100)
    lib = H.synth_library(cfg)
    ctx = H.Context(0)
    ctx.set_library(lib)
    ctx.reserve(h * w, b)
    ps = Prescreen(ctx, b, seed=7)
    t0 = time.perf_counter()
    for f in range(train_frames):
        cube = torch.from_numpy(H.synth_frame(cfg, f, threads=16)).cuda()
        ps.add_training_frame(cube, h, w)
        del cube
    history = ps.train(epochs=60, batch=1024)
    cut = ps.calibrate()
    train_s = time.perf_counter() - t0
    log(f"  learned: trained in {train_s:.1f} s, final loss {history[-1]:.4f}, cut {cut:.3f}")
    evals = []
    cube = None
    for f in range(eval_frames):
        cube = torch.from_numpy(H.synth_frame(cfg, 1000 + f, threads=16)).cuda()
        evals.append(ps.evaluate(cube, h, w))
    # timing on the last held-out frame (statistics are built for it)
    ms = {"plain_score": time_on_stream(torch, lambda: ps.plain(cube, h, w), repeats),
          "learned_score": time_on_stream(torch, lambda: ps.screen(cube, h, w), repeats),
          "learned_score_two_pass": time_on_stream(torch, lambda: ps.screen_two_pass(cube, h, w), repeats),
          "stats_plain": time_on_stream(torch, lambda: (ctx.stats_accumulate(cube, h, w, b), ctx.stats_build(LOADING)), repeats),
          "stats_mixed": time_on_stream(torch, lambda: (ctx.stats_accumulate(cube, h, w, b, precision=H.MIXED), ctx.stats_build(LOADING)), repeats)}
    m = ps._buffers(h * w)
    logit = m["logit"]
    ms["infer"] = time_on_stream(torch, lambda: ctx.mlp_infer(cube, h, w, b, logit, center=True), repeats)
    # SAM and CEM alone on the same 2048x2048x113 frame (8 targets)
    s_map, t_map = m["ace"], m["target"]
    ms["sam"] = time_on_stream(torch, lambda: ctx.score(H.Detector.SAM, cube, h, w, b, s_map, t_map), repeats)
    ms["cem"] = time_on_stream(torch, lambda: ctx.score(H.Detector.CEM, cube, h, w, b, s_map, t_map), repeats)
    ctx.close()
    log("  learned: " + ", ".join(f"{k} {v * 1e3:.1f} us" for k, v in ms.items()))
    rec = {d: min(e[d]["recall"] for e in evals) for d in ("rx", "ace")}
    pix = h * w
    return {
        "workload": f"RX+ACE, {h}x{w}x{b}, {t} targets, 1% anomalies; {train_frames} training frames, "
                    f"{eval_frames} held-out frames",
        "training_seconds": round(train_s, 2),
        "cut": round(cut, 4),
        "min_recall": rec,
        "false_skips": {d: int(sum(e[d]["false_skips"] for e in evals)) for d in ("rx", "ace")},
        "plain_detections": {d: int(sum(e[d]["plain"] for e in evals)) for d in ("rx", "ace")},
        "candidate_fraction": round(float(np.mean([e["candidate_fraction"] for e in evals])), 5),
        "us": {k: round(v * 1e3, 1) for k, v in ms.items()},
        "score_mpix_s": {"plain": round(pix / (ms["plain_score"] / 1e3) / 1e6, 1),
                         "learned": round(pix / (ms["learned_score"] / 1e3) / 1e6, 1)},
        "speedup_score": round(ms["plain_score"] / ms["learned_score"], 3),
        "speedup_path_with_statistics": round((ms["stats_plain"] + ms["plain_score"]) /
                                              (ms["stats_mixed"] + ms["learned_score"]), 3),
        "infer_hbm_gb_s": round(b * 4 * pix / (ms["infer"] / 1e3) / 1e9, 1),
        "sam_cem_2048": {k: {"us": round(ms[k] * 1e3, 1),
                             "hbm_gb_s": round((b * 4 + 8) * pix / (ms[k] / 1e3) / 1e9, 1),
                             "hbm_frac": round((b * 4 + 8) * pix / (ms[k] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
                         for k in ("sam", "cem")},
    }


def batched_form(H, torch, repeats, seed=20261203 + 4000):
    """64 uneven scenes in one batch: heights U{128..2048}, widths U{128..1024},
    bands from {16, 113, 224}, libraries of U{1..32} targets; 16 of them are
    4x4 snapshot-mosaic frames of 2048x1088 (u16) demosaiced on the device to
    512x272x16. All four detectors run. Timed: the batch's persistent
    launches (one per shape class, shapes from the autotuner) against the
    same scoring issued scene by scene through the single-scene entry points."""
    rng = np.random.default_rng(seed)
    f32 = dict(dtype=torch.float32, device="cuda")
    i32 = dict(dtype=torch.int32, device="cuda")
    scenes = []
    total_px = 0
    ctx0 = H.Context(0)
    demosaic_us = 0.0
    for i in range(64):
        t = int(rng.integers(1, 33))
        if i % 4 == 3:   # a snapshot-mosaic frame
            h, w, b = 272, 512, 16
        else:
            h, w = int(rng.integers(128, 2049)), int(rng.integers(128, 1025))
            b = int(rng.choice([16, 113, 224]))
        cfg = H.synth_config(h, w, b, min(t, 8), seed=seed + i, anomaly_fraction=0.005)
        lib = H.synth_library(H.synth_config(h, w, b, t, seed=seed + i))
        stride = H.plane_stride(h, w)
        if i % 4 == 3:
            host = H.synth_frame(cfg, 0, threads=16)[:, :h * w].reshape(b, h, w)
            raw = np.zeros((h * 4, w * 4), np.uint16)
            for band in range(b):
                oy, ox = divmod(band, 4)
                raw[oy::4, ox::4] = np.clip(np.round(host[band] * 65535.0), 0, 65535).astype(np.uint16)
            raw_d = torch.from_numpy(raw.view(np.int16)).cuda()
            cube = torch.zeros((b, stride), **f32)
            demosaic_us += time_on_stream(torch, lambda: H.demosaic(ctx0, raw_d, H.RAW_U16, h * 4, w * 4, 4, b, cube), 3) * 1e3
            del raw_d
        else:
            cube = torch.from_numpy(H.synth_frame(cfg, 0, threads=16)).cuda()
        p = h * w
        total_px += p
        scenes.append(dict(cube=cube, height=h, width=w, library=lib,
                           sam_score=torch.empty(p, **f32), sam_target=torch.empty(p, **i32),
                           cem_score=torch.empty(p, **f32), cem_target=torch.empty(p, **i32),
                           rx=torch.empty(p, **f32), ace_score=torch.empty(p, **f32),
                           ace_target=torch.empty(p, **i32)))
    batch = H.Batch(0)
    batch.set_scenes(scenes)
    batch.statistics(LOADING)
    batch.check()
    report = batch.autotune()
    log("  batch autotuner:\n    " + report.strip().replace("\n", "\n    "))
    ms_batch = time_on_stream(torch, batch.score, repeats)
    ctxs = [batch.context(i) for i in range(len(scenes))]

    def per_scene():
        for c, sc in zip(ctxs, scenes):
            b = sc["cube"].shape[0]
            c.score_sam_cem(sc["cube"], sc["height"], sc["width"], b, sc["sam_score"], sc["sam_target"],
                            sc["cem_score"], sc["cem_target"])
            c.score(H.Detector.ACE, sc["cube"], sc["height"], sc["width"], b, sc["ace_score"],
                    sc["ace_target"], sc["rx"])

    ms_single = time_on_stream(torch, per_scene, repeats)
    ms_stats = time_on_stream(torch, lambda: batch.statistics(LOADING), repeats)
    classes = report.strip().count("\n") + 1
    ctx0.close()
    batch.close()
    log(f"  batched: {ms_batch * 1e3:.1f} us for {total_px / 1e6:.1f} Mpix, scene by scene {ms_single * 1e3:.1f} us")
    return {
        "workload": "64 uneven scenes (H 128-2048, W 128-1024, bands 16/113/224, 1-32 targets; 16 demosaiced "
                    "2048x1088 4x4 mosaics), SAM+CEM+RX+ACE",
        "pixels": total_px,
        "launches": classes,
        "batched_us": round(ms_batch * 1e3, 1),
        "scene_by_scene_us": round(ms_single * 1e3, 1),
        "batched_mpix_s": round(total_px / (ms_batch / 1e3) / 1e6, 1),
        "speedup_vs_scene_by_scene": round(ms_single / ms_batch, 3),
        "statistics_us": round(ms_stats * 1e3, 1),
        "demosaic_us_total": round(demosaic_us, 1),
        "autotuner": report.strip().split("\n"),
    }


def main():
    args = parse()
    if args.copy_calibrate:
        copy_calibrate(args.copy_calibrate)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has a single MI355X, so "
              f"--gpus {args.gpus} is not supported (the path would shard whole frames across "
              f"GPUs with no exchange).", file=sys.stderr)
        sys.exit(2)
    import torch
    import hsi_mi355x as H

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    pixels = H_ * W_
    stride = H.plane_stride(H_, W_)
    cfg = H.synth_config(H_, W_, B_, T_, drift=0.2, drift_period=300)
    library = H.synth_library(cfg)
    log(f"generating {args.frames} frames {H_}x{W_}x{B_} on the host")
    frames = []
    host0 = None
    for f in range(args.frames):
        host = H.synth_frame(cfg, f, threads=16)
        if f == 0:
            host0 = host
        frames.append(torch.from_numpy(host).cuda())
    torch.cuda.synchronize()
    log("frames resident in HBM")

    ctx = H.Context(0)
    ctx.set_library(library)
    ctx.reserve(pixels, B_)
    # SAM runs on its own stream; it gets its own context because a context's
    # scratch must not be shared by launches that overlap.
    ctx_sam = H.Context(0)
    ctx_sam.set_library(library)
    ctx_sam.reserve(pixels, B_)
    side = torch.cuda.Stream()
    main_stream = torch.cuda.current_stream()
    f32 = dict(dtype=torch.float32, device="cuda")
    i32 = dict(dtype=torch.int32, device="cuda")
    sam_s, sam_t = torch.empty(pixels, **f32), torch.empty(pixels, **i32)
    cem_s, cem_t = torch.empty(pixels, **f32), torch.empty(pixels, **i32)
    cap = 4096
    sam_d, cem_d = torch.empty((cap, 4), **i32), torch.empty((cap, 4), **i32)
    sam_n, cem_n = torch.empty(1, **i32), torch.empty(1, **i32)
    p_sam = H.DetectionParams(SAM_THRESHOLD, 2, cap, False)
    p_cem = H.DetectionParams(CEM_THRESHOLD, 2, cap, True)

    def stats(cube):
        ctx.stats_accumulate(cube, H_, W_, B_)

    def build():
        ctx.stats_build(LOADING)

    # Frames in flight, as the reference's Pipeline keeps several (its
    # --streams): slot k owns a context and a stream for the statistics + CEM
    # chain, another pair for SAM, and its own score and detection buffers.
    # Frames go to slots round robin; a slot's streams keep its frames in
    # order, and different slots' frames overlap (one slot's single-CU build
    # runs beside another's moments). With one slot, SAM of a frame starts
    # only after the previous frame finished, so frames do not overlap at all.
    inflight = max(1, args.in_flight)
    slots = []
    for k in range(inflight):
        if k == 0:
            c, cs, a, b_ = ctx, ctx_sam, main_stream, side
            bufs = (sam_s, sam_t, cem_s, cem_t, sam_d, cem_d, sam_n, cem_n)
        else:
            c, cs = H.Context(0), H.Context(0)
            for x in (c, cs):
                x.set_library(library)
                x.reserve(pixels, B_)
            a, b_ = torch.cuda.Stream(), torch.cuda.Stream()
            bufs = (torch.empty(pixels, **f32), torch.empty(pixels, **i32), torch.empty(pixels, **f32),
                    torch.empty(pixels, **i32), torch.empty((cap, 4), **i32), torch.empty((cap, 4), **i32),
                    torch.empty(1, **i32), torch.empty(1, **i32))
        slots.append((c, cs, a, b_, bufs))

    def step(i):
        cube = frames[i % len(frames)]
        c, cs, a, b_, (ss, st, cms, cmt, sd, cd, sn, cn) = slots[i % inflight]
        if args.separate_sam:
            if inflight == 1:
                b_.wait_stream(a)
            cs.score(H.Detector.SAM, cube, H_, W_, B_, ss, st, stream=b_)
            cs.detect(ss, st, H_, W_, p_sam, sd, sn, stream=b_)
        c.stats_accumulate(cube, H_, W_, B_, stream=a)
        c.stats_build(LOADING, stream=a)
        if args.separate_sam:
            c.score(H.Detector.CEM, cube, H_, W_, B_, cms, cmt, stream=a)
        else:
            # SAM and CEM from one read of the cube once the statistics exist
            c.score_sam_cem(cube, H_, W_, B_, ss, st, cms, cmt, stream=a)
            c.detect(ss, st, H_, W_, p_sam, sd, sn, stream=a)
        c.detect(cms, cmt, H_, W_, p_cem, cd, cn, stream=a)
        if args.separate_sam and inflight == 1:
            a.wait_stream(b_)

    step(0)
    ctx.stats_check()
    torch.cuda.synchronize()
    log("warm-up")
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the slots the timed steps used (fewer steps than slots leave some idle)
    for c, *_ in slots[:min(inflight, args.steps)]:
        c.stats_check()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = pixels * args.steps / (t1 - t0) / 1e6
    log(f"step {ms_step:.3f} ms, {value:.1f} Mpix/s")

    # Per-kernel timing on the launch stream.
    c0 = frames[0]
    r = args.repeats
    ms = {
        "moments": time_on_stream(torch, lambda: stats(c0), r),
        "moments_mixed": time_on_stream(torch, lambda: ctx.stats_accumulate(c0, H_, W_, B_, precision=H.MIXED), r),
        "build": time_on_stream(torch, build, r),
        "sam_cem_fused": time_on_stream(torch, lambda: ctx.score_sam_cem(c0, H_, W_, B_, sam_s, sam_t, cem_s, cem_t), r),
        "sam": time_on_stream(torch, lambda: ctx.score(H.Detector.SAM, c0, H_, W_, B_, sam_s, sam_t), r),
        "cem": time_on_stream(torch, lambda: ctx.score(H.Detector.CEM, c0, H_, W_, B_, cem_s, cem_t), r),
        "detect": time_on_stream(torch, lambda: ctx.detect(sam_s, sam_t, H_, W_, p_sam, sam_d, sam_n), r),
        "rx": time_on_stream(torch, lambda: ctx.score(H.Detector.RX, c0, H_, W_, B_, cem_s), r),
        "ace": time_on_stream(torch, lambda: ctx.score(H.Detector.ACE, c0, H_, W_, B_, cem_s, cem_t, sam_s), r),
        "rx_mixed": time_on_stream(torch, lambda: ctx.score(H.Detector.RX, c0, H_, W_, B_, cem_s, precision=H.MIXED), r),
        "ace_mixed": time_on_stream(torch, lambda: ctx.score(H.Detector.ACE, c0, H_, W_, B_, cem_s, cem_t, sam_s, precision=H.MIXED), r),
    }
    for k, v in ms.items():
        log(f"  {k:14s} {v * 1e3:9.1f} us")

    cube_bytes = B_ * 4
    # Algorithmic counts per pixel (bytes read + written; FP64 flops).
    bytes_px = {"sam": cube_bytes + 8, "cem": cube_bytes + 8, "sam_cem_fused": cube_bytes + 16,
                "moments": cube_bytes, "rx": cube_bytes + 4, "ace": cube_bytes + 12, "detect": 8}
    flops_px = {"sam": 2 * B_ * (T_ + 1), "cem": 2 * B_ * T_, "sam_cem_fused": 2 * B_ * (2 * T_ + 1),
                "moments": B_ * (B_ + 1), "rx": B_ * B_ + 4 * B_,
                "ace": B_ * B_ + 4 * B_ + 2 * B_ * T_}
    detectors = {}
    bytes_px["moments_mixed"] = cube_bytes
    flops_px["moments_mixed"] = flops_px["moments"]
    for k in ("sam", "cem", "sam_cem_fused", "rx", "ace", "moments", "moments_mixed"):
        sec = ms[k] / 1e3
        gbs = bytes_px[k] * pixels / sec / 1e9
        tfl = flops_px[k] * pixels / sec / 1e12
        detectors[k] = {"mpix_s": round(pixels / sec / 1e6, 1), "us": round(ms[k] * 1e3, 1),
                        "hbm_gb_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                        "fp64_tflops": round(tfl, 2), "fp64_frac": round(tfl / FP64_PEAK_TFLOPS, 4)}

    # Dominant kernel of the step: moments (FP64 matrix cores) vs the fused
    # SAM+CEM stream (HBM).
    step_kernels = {"moments": ms["moments"], "sam_cem_fused": ms["sam_cem_fused"],
                    "build": ms["build"], "detect": 2 * ms["detect"]}
    dom = max(step_kernels, key=step_kernels.get)
    if dom == "moments":
        roof = {"kernel": "moments_f64_kernel", "bound": "mfma",
                "achieved": round(flops_px["moments"] * pixels / (ms["moments"] / 1e3) / 1e12, 3),
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "per_unit": f"{flops_px['moments']} FP64 flop/pixel (B(B+1)), {pixels} pixels/launch"}
    elif dom == "sam_cem_fused":
        roof = {"kernel": "sam_cem_fused_kernel", "bound": "hbm",
                "achieved": round(bytes_px["sam_cem_fused"] * pixels / (ms["sam_cem_fused"] / 1e3) / 1e9, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "per_unit": f"{bytes_px['sam_cem_fused']} B/pixel, {pixels} pixels/launch"}
    else:
        roof = {"kernel": dom, "bound": "latency", "achieved": None, "peak": None, "unit": None}
    if roof.get("achieved") is not None:
        roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    roof["traffic"] = round(MEASURED_TRAFFIC_BYTES[roof["kernel"]]) if roof["kernel"] in MEASURED_TRAFFIC_BYTES else None

    prec = None
    if not args.no_precision:
        log("precision form (224 bands)")
        prec = precision_form(H, torch, max(3, r // 4))

    batched = None
    if not args.no_batch:
        log("batched uneven scenes (64)")
        batched = batched_form(H, torch, max(3, r // 4))

    learned = None
    if not args.no_learned:
        log("learned pre-screen (2048x2048x113)")
        learned = learned_form(H, torch, max(3, r // 4))

    cpu = None
    if not args.no_cpu:
        log("CPU baseline")
        cpu = cpu_baseline(host0, library, args.cpu_rows)
        log(f"  CPU {cpu['value']} Mpix/s ({cpu['kind']}, {cpu['seconds']} s)")

    line = {
        "metric": "Mpix/s per detector vs HBM/MFMA roofline; mixed+learned vs plain fp32",
        "value": round(value, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 in, f64 accumulate",
        "data": "synthetic (seeded generator, frames resident in HBM)",
        "config": {"workload": "plain FP32 video frame: moments+build+fused SAM/CEM+detect x2, frames in flight" if not args.separate_sam else "plain FP32 video frame: moments+build+CEM+detect || SAM+detect, frames in flight",
                   "height": H_, "width": W_, "bands": B_, "targets": T_,
                   "resident_frames": args.frames, "frames_in_flight": max(1, args.in_flight), "loading": LOADING},
        "roofline": roof,
        "cpu_baseline": cpu,
        "detectors": detectors,
        "precision_form": prec,
        "learned_form": learned,
        "batched_form": batched,
        "stage_us": {k: round(v * 1e3, 1) for k, v in ms.items()},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
