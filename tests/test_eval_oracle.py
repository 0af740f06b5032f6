"""Pins the evaluation oracle (oracle/eval_ref.py) by mathematics alone: the
rank-sum AUC equals the fraction of (positive, negative) pairs ordered
correctly, ties counting half (brute force), scipy's Mann-Whitney U / (P N),
and the closed forms for perfect separation (1), reversal (0) and all ties
(1/2); the operating-point counts partition the annotated pixels."""
import numpy as np
import pytest

from oracle import eval_ref


def brute_auc(scores, labels):
    pos = scores[labels]
    neg = scores[~labels]
    gt = (pos[:, None] > neg[None, :]).sum()
    eq = (pos[:, None] == neg[None, :]).sum()
    return (gt + 0.5 * eq) / (pos.size * neg.size)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_auc_matches_pair_counting_with_ties(seed):
    rng = np.random.default_rng(seed)
    scores = np.round(rng.normal(0, 1, 400), 1)     # many ties
    labels = rng.random(400) < 0.3
    assert abs(eval_ref.roc_auc(scores, labels) - brute_auc(scores, labels)) < 1e-12


def test_auc_matches_scipy_mann_whitney():
    from scipy.stats import mannwhitneyu
    rng = np.random.default_rng(7)
    scores = np.round(rng.normal(0, 1, 3000), 2)
    labels = rng.random(3000) < 0.1
    u = mannwhitneyu(scores[labels], scores[~labels], alternative="two-sided").statistic
    assert abs(eval_ref.roc_auc(scores, labels) - u / (labels.sum() * (~labels).sum())) < 1e-12


def test_auc_closed_forms():
    labels = np.array([0, 0, 1, 1, 0, 1], bool)
    assert eval_ref.roc_auc(np.where(labels, 2.0, 1.0), labels) == 1.0
    assert eval_ref.roc_auc(np.where(labels, 1.0, 2.0), labels) == 0.0
    assert eval_ref.roc_auc(np.full(6, 3.0), labels) == 0.5
    assert np.isnan(eval_ref.roc_auc(np.arange(3.0), np.zeros(3, bool)))


def test_operating_point_partitions_annotated_pixels():
    rng = np.random.default_rng(3)
    truth = rng.integers(0, 12, 5000).astype(np.uint8)      # 9..11 unannotated
    angle = rng.uniform(0, 0.3, 5000).astype(np.float32)
    r = eval_ref.evaluate(angle, truth, target=1, max_label=8, threshold=0.1)
    valid = truth <= 8
    assert r["tp"] + r["fp"] + r["fn"] + r["tn"] == valid.sum()
    assert r["tp"] + r["fn"] == (truth == 1).sum() == r["positives"]
    assert r["tp"] == ((truth == 1) & (angle <= 0.1)).sum()
    assert sum(c[0] for c in r["classes"].values()) == valid.sum()
