"""The learned pre-screen at its workload size: 2048 x 2048 px x 113 bands,
8 targets, 1% anomalies. The network is trained on the GPU from the plain
path's labels of three frames and judged on a held-out frame:
  - over the full frame, recall of the plain RX and ACE detection lists is
    >= 0.999, and the false skips are counted;
  - on a 256-row stripe, the learned path's lists are compared with
    detect_cpu on the oracle's own scores (oracle statistics from the whole
    frame). Every difference must be explained: a pixel within 1e-5 of the
    threshold or of an NMS neighbour, a counted false skip, or a pixel next
    to one (a skipped neighbour no longer suppresses it).
The 20-training / 100-held-out protocol of the workload runs as a bench leg
(bench.py --learned-frames 20,100), not as a test."""
import os

import numpy as np
import pytest

from gpu_util import near_cut

pytestmark = pytest.mark.gpu

SIZE, B, T = 2048, 113, 8
STRIPE, NMS = 256, 2
THREADS = min(16, os.cpu_count() or 1)


def frame_cfg(H):
    return H.synth_config(SIZE, SIZE, B, T, seed=20261203 + 5000, anomaly_fraction=0.01, drift=0.2,
                          drift_period=100)


def as_set(dets, rows=None):
    pts = {(int(x), int(y)) for x, y in dets}
    return pts if rows is None else {(x, y) for x, y in pts if y < rows}


def test_learned_path_at_2048_against_plain_and_oracle(oracle_lib):
    import torch
    import hsi_mi355x as H
    from hsi_mi355x.prescreen import Prescreen
    cfg = frame_cfg(H)
    lib = H.synth_library(cfg)
    ctx = H.Context(0)
    ctx.set_library(lib)
    ctx.reserve(SIZE * SIZE, B)
    ps = Prescreen(ctx, B, seed=7)
    for f in range(3):
        print(f"training frame {f}", flush=True)
        cube = torch.from_numpy(H.synth_frame(cfg, f, threads=THREADS)).cuda()
        ps.add_training_frame(cube, SIZE, SIZE)
        del cube
    ps.train(epochs=60, batch=1024)
    cut = ps.calibrate()
    print(f"held-out frame, cut {cut:.3f}", flush=True)
    host = H.synth_frame(cfg, 1000, threads=THREADS)
    cube = torch.from_numpy(host).cuda()
    ps.statistics(cube, SIZE, SIZE)
    plain_rx, plain_ace = ps.detections(ps.plain(cube, SIZE, SIZE), SIZE, SIZE)
    m = ps.screen(cube, SIZE, SIZE)
    torch.cuda.synchronize()
    flagged = int(m["count"].item())
    learned_rx, learned_ace = ps.detections(m, SIZE, SIZE)
    ctx.close()
    skips = {}
    for name, want, got in (("rx", plain_rx, learned_rx), ("ace", plain_ace, learned_ace)):
        skips[name] = want - got
        recall = 1 - len(skips[name]) / max(len(want), 1)
        print(f"{name}: {len(want)} plain, {len(got)} learned, recall {recall:.6f}, "
              f"false skips {len(skips[name])}; candidates {flagged} ({flagged / SIZE ** 2:.4f})")
        assert len(want) > 1000
        assert recall >= 0.999

    print("oracle statistics and stripe scores", flush=True)
    p = SIZE * SIZE
    full = np.ascontiguousarray(host[:, :p]).reshape(B, SIZE, SIZE)
    count, s1, s2 = oracle_lib.moments(full, pixels=p, threads=THREADS)
    st = oracle_lib.build_statistics(count, s1, s2, lib, 1e-6)
    assert st.valid, st.error
    stripe = np.ascontiguousarray(full[:, :STRIPE, :])
    ace, ace_t, rx = oracle_lib.ace_best(stripe, st, threads=THREADS)
    rows = STRIPE - NMS  # detections in the last NMS rows depend on rows below the stripe
    for name, score, target, thr, got in (("rx", rx, None, ps.rx_thr, learned_rx),
                                          ("ace", ace, ace_t, ps.ace_thr, learned_ace)):
        want_d, _ = oracle_lib.detect(score, target, STRIPE, SIZE, thr, NMS, True)
        want = as_set(zip(want_d["x"], want_d["y"]), rows)
        got_s = {d for d in got if d[1] < rows}
        near = near_cut(score, STRIPE, SIZE, thr, NMS, 1e-5).reshape(STRIPE, SIZE)
        skipped = skips[name]

        def beside_skip(x, y):
            return any((x + dx, y + dy) in skipped for dx in range(-NMS, NMS + 1) for dy in range(-NMS, NMS + 1))

        missing = [d for d in want - got_s if not near[d[1], d[0]] and d not in skipped]
        extra = [d for d in got_s - want if not near[d[1], d[0]] and not beside_skip(*d)]
        print(f"stripe {name}: oracle {len(want)}, learned {len(got_s)}, differences explained "
              f"{len(want ^ got_s) - len(missing) - len(extra)}")
        assert len(want) > 50
        assert not missing, missing[:10]
        assert not extra, extra[:10]
