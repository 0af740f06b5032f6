"""hsi_evaluate on the GPU against the evaluation oracle (oracle/eval_ref.py,
pinned by pair counting and scipy): AUC to 1e-12 (both are exact rationals
rounded once), the operating-point counts and per-class counts and minima
exactly, per-class means to 1e-12 relative; on SAM-like angles with heavy
ties, unannotated labels, NaN-free maps of several sizes, and a higher-is-
better map. Also the CLI on dumped files."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG, REPO
from oracle import eval_ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import torch
    import hsi_mi355x
    assert torch.cuda.is_available()
    return hsi_mi355x


def case(n, seed, ties):
    rng = np.random.default_rng(seed)
    truth = rng.choice(np.arange(12, dtype=np.uint8), n, p=[0.6, 0.07] + [0.03] * 7 + [0.04] * 3)
    angle = rng.uniform(0.0, 0.4, n) - 0.15 * (truth == 1)
    angle = np.clip(angle, 0.0, None).astype(np.float32)
    if ties:
        angle = (np.round(angle * 50) / 50).astype(np.float32)
    return angle, truth


@pytest.mark.parametrize("n,seed,ties,higher", [(1000, 1, True, False), (65536 + 37, 2, False, False),
                                                (300_000, 3, True, False), (4096, 4, True, True)])
def test_evaluate_matches_oracle(H, n, seed, ties, higher):
    import torch
    angle, truth = case(n, seed, ties)
    score = -angle if higher else angle
    thr = -0.1 if higher else 0.1
    ctx = H.Context(0)
    print(f"evaluate n={n} ties={ties} higher={higher}", flush=True)
    got = ctx.evaluate(torch.from_numpy(score).cuda(), torch.from_numpy(truth).cuda(), 1, 8, higher, thr)
    ctx.close()
    want = eval_ref.evaluate(score, truth, 1, 8, higher, thr)
    assert abs(got["auc"] - want["auc"]) <= 1e-12, (got["auc"], want["auc"])
    for k in ("positives", "negatives", "tp", "fp", "fn", "tn"):
        assert got[k] == want[k], k
    assert got["classes"].keys() == want["classes"].keys()
    for c, (cnt, mean, mn) in want["classes"].items():
        g = got["classes"][c]
        assert g[0] == cnt and g[2] == mn
        assert abs(g[1] - mean) <= 1e-12 * max(1.0, abs(mean))


def test_evaluate_cli_report(H, tmp_path):
    angle, truth = case(520 * 696, 5, True)
    angle.astype("<f4").tofile(tmp_path / "map.f32")
    truth.tofile(tmp_path / "gt.u8")
    out = subprocess.run([sys.executable, "-m", "hsi_mi355x.evaluate", str(tmp_path / "map.f32"),
                          str(tmp_path / "gt.u8"), "--width=696", "--height=520"],
                         capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, PYTHONPATH=os.pathsep.join([PKG, REPO])))
    assert out.returncode == 0, out.stderr
    want = eval_ref.evaluate(angle, truth)
    assert f"ROC AUC           {want['auc']:.4f}" in out.stdout
    assert f"tp {want['tp']}  fp {want['fp']}  fn {want['fn']}  tn {want['tn']}" in out.stdout
