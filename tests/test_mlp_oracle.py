"""Pins the pre-screen's CPU oracle (oracle/mlp_ref.py) by mathematics alone:
central finite differences of its loss against its analytic gradient, the
bf16 rounding helper against hand-worked bit patterns, and Adam's first step
against its closed form (a step of exactly lr * sign(g))."""
import numpy as np
import pytest

from oracle import mlp_ref


def tiny_problem(bands=7, n=9, seed=3):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.4, mlp_ref.param_count(bands))
    x = rng.normal(0, 1.0, (bands, n))
    y = (rng.random(n) < 0.4).astype(np.float64)
    return w, x, y


@pytest.mark.parametrize("pos_weight", [1.0, 7.5])
def test_gradient_matches_central_differences(pos_weight):
    bands = 7
    w, x, y = tiny_problem(bands)
    loss, g = mlp_ref.loss_and_grad(w, x, y, pos_weight, bands)
    eps = 1e-6
    rng = np.random.default_rng(11)
    # every bias and w3/b3 entry, plus a random sample of W1 and W2 entries
    P = mlp_ref.param_count(bands)
    idx = set(range(64 * bands, P))
    idx.update(rng.choice(64 * bands, 60, replace=False).tolist())
    checked = 0
    for i in sorted(idx):
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        num = (mlp_ref.loss_and_grad(wp, x, y, pos_weight, bands)[0] -
               mlp_ref.loss_and_grad(wm, x, y, pos_weight, bands)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])
        checked += 1
    assert checked > 2000 and np.isfinite(loss)


def test_loss_is_weighted_binary_cross_entropy():
    bands = 5
    w, x, y = tiny_problem(bands, n=4)
    z = mlp_ref.forward(w, x, bands)[0]
    p = 1 / (1 + np.exp(-z))
    want = np.mean(-(3.0 * y * np.log(p) + (1 - y) * np.log(1 - p)))
    assert abs(mlp_ref.loss_and_grad(w, x, y, 3.0, bands)[0] - want) < 1e-12


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0, 1.0 + 2 ** -8 + 2 ** -9],
                    np.float32)
    got = mlp_ref.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0); 1 + 3*2^-9 rounds up
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0078125
    assert got[5] == 1.0078125
    bits = got.astype(np.float32).view(np.uint32)
    assert np.all(bits & 0xFFFF == 0)


def test_adam_first_step_is_lr_times_sign():
    g = np.array([0.3, -2.0, 1e-3])
    w = np.zeros(3)
    w1, m, v = mlp_ref.adam_step(w, np.zeros(3), np.zeros(3), g, 1, 0.01, 0.9, 0.999, 0.0)
    assert np.allclose(w1, -0.01 * np.sign(g))
