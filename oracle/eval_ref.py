"""CPU oracle for the offline evaluation (test infrastructure only).

A numpy restatement of the reference's scripts/eval_detection.py: roc_auc
(:40-66) - Mann-Whitney AUC from the rank sum, ties given their average rank
- the operating point at a threshold (:96-108) and the per-class count, mean
and minimum of the score (:117-132). Pinned by tests/test_eval_oracle.py
against brute-force pair counting, scipy's Mann-Whitney U statistic and
closed forms (perfect separation, reversal, all ties).
"""
import numpy as np


def roc_auc(scores, labels):
    """AUC of `scores` (higher = more positive) for boolean `labels`, via the
    rank sum with tie-averaged ranks (eval_detection.py:40-66)."""
    scores = np.asarray(scores, np.float64)
    labels = np.asarray(labels).astype(bool)
    positives = int(labels.sum())
    negatives = labels.size - positives
    if positives == 0 or negatives == 0:
        return float("nan")
    order = np.argsort(scores, kind="mergesort")
    ranks = np.empty(scores.size, np.float64)
    ranks[order] = np.arange(1, scores.size + 1, dtype=np.float64)
    sorted_scores = scores[order]
    # runs of equal scores share the mean of their ranks
    starts = np.flatnonzero(np.r_[True, sorted_scores[1:] != sorted_scores[:-1]])
    ends = np.r_[starts[1:], scores.size]
    for s, e in zip(starts, ends):
        if e - s > 1:
            ranks[order[s:e]] = (s + 1 + e) / 2.0
    rank_sum = ranks[labels].sum()
    return (rank_sum - positives * (positives + 1) / 2.0) / (positives * negatives)


def evaluate(score, truth, target=1, max_label=8, higher_is_better=False, threshold=0.10):
    """The eval_detection.py report as numbers: annotated pixels are
    truth <= max_label, positives truth == target; AUC on the score (negated
    when lower is better); counts at the threshold; per-class stats."""
    score = np.asarray(score, np.float32)
    truth = np.asarray(truth, np.uint8)
    valid = truth <= max_label
    lab = truth[valid] == target
    s = score[valid].astype(np.float64)
    keys = s if higher_is_better else -s
    pred = (score[valid] >= threshold) if higher_is_better else (score[valid] <= threshold)
    out = {"auc": roc_auc(keys, lab), "positives": int(lab.sum()), "negatives": int((~lab).sum()),
           "tp": int((pred & lab).sum()), "fp": int((pred & ~lab).sum()),
           "fn": int((~pred & lab).sum()), "tn": int((~pred & ~lab).sum()), "classes": {}}
    for c in np.unique(truth[valid]):
        sel = s[truth[valid] == c]
        out["classes"][int(c)] = (int(sel.size), float(sel.mean()), float(sel.min()))
    return out
