"""The learned pre-screen end to end: trained on the GPU from plain RX/ACE
labels of seeded frames with planted targets and 1% anomalies, then judged on
held-out frames against the plain full-frame detectors (recall of their
detection lists), and against the same procedure with an untrained network
(what training gains: the share of pixels that still need exact scoring)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H_, W_, B_, T_ = 256, 256, 113, 8


def frames(H, idx):
    import torch
    cfg = H.synth_config(H_, W_, B_, T_, seed=777, anomaly_fraction=0.01, drift=0.2, drift_period=50)
    lib = H.synth_library(cfg)
    return lib, [torch.from_numpy(H.synth_frame(cfg, i, threads=8)).cuda() for i in idx]


def run_prescreen(H, train_epochs):
    from hsi_mi355x.prescreen import Prescreen
    lib, train = frames(H, [0, 1, 2, 3])
    _, held = frames(H, [10, 11, 12])
    ctx = H.Context(0)
    ctx.set_library(lib)
    ps = Prescreen(ctx, B_, seed=3)
    for i, cube in enumerate(train):
        npos, nneg = ps.add_training_frame(cube, H_, W_)
        print(f"training frame {i}: {npos} positives, {nneg} negatives", flush=True)
    hist = ps.train(epochs=train_epochs)
    cut = ps.calibrate()
    results = [ps.evaluate(c, H_, W_) for c in held]
    ctx.close()
    return hist, cut, results


def test_trained_prescreen_keeps_plain_detections_and_skips_most_pixels():
    import hsi_mi355x as H
    hist, cut, results = run_prescreen(H, 60)
    print("loss every 10 epochs", [round(v, 4) for v in hist[::10]], "cut", round(cut, 3))
    for r in results:
        print(r)
        for det in ("rx", "ace"):
            assert r[det]["plain"] > 20
            assert r[det]["recall"] >= 0.999, (det, r)
        assert r["candidate_fraction"] < 0.10
    _, ucut, untrained = run_prescreen(H, 0)
    fr_u = np.mean([r["candidate_fraction"] for r in untrained])
    fr_t = np.mean([r["candidate_fraction"] for r in results])
    print(f"candidate fraction: trained {fr_t:.4f}, untrained {fr_u:.4f}")
    assert fr_t < 0.5 * fr_u
