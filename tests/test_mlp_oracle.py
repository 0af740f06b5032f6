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


def test_forward_bf16_equals_forward_on_bf16_exact_inputs():
    """Weights, inputs, biases and layer-1 activations that bfloat16 holds
    exactly (multiples of 1/16 below 8): every rounding step is the identity,
    so the bf16 forward pass must equal the float64 one."""
    bands, n = 3, 40
    rng = np.random.default_rng(5)
    P = mlp_ref.param_count(bands)
    w = rng.integers(-4, 5, P) / 4.0
    W1, b1, W2, b2, w3, b3 = mlp_ref.unpack(w, bands)
    b1[:] = rng.integers(-8, 9, b1.shape) / 16.0
    w = mlp_ref.pack(W1, b1, W2, b2, w3, b3)
    x = rng.integers(-4, 5, (bands, n)) / 4.0
    z, h1, _ = mlp_ref.forward(w, x, bands)
    assert np.array_equal(mlp_ref.bf16(h1), h1)
    assert np.abs(h1).max() > 0
    assert np.array_equal(mlp_ref.forward_bf16(w, x, bands), z)


def test_forward_bf16_hand_worked():
    """One live unit per layer, worked by hand. bf16 keeps 8 significant bits:
    bf16(1.01) = 1.0078125 (129/128), bf16(0.3) = 0.30078125 (154/512),
    bf16(0.51) = 0.51171875 (131/256).
    Sample 0, x = (3, 0.3): W1.x = 1.0078125*3 + 2*0.30078125 = 3.625;
      + b1 = 3.725 -> bf16 3.71875 (238/64); W2 h1 = 0.51171875*3.71875 =
      1.9029541015625; + b2 = 2.1529541015625; logit = 2*that - 1 = 3.305908203125.
    Sample 1, x = (-1, 0): W1.x + b1 = -0.9078125 -> relu 0; h2 = b2 = 0.25;
      logit = -0.5."""
    bands = 2
    w = np.zeros(mlp_ref.param_count(bands))
    W1, b1, W2, b2, w3, b3 = mlp_ref.unpack(w, bands)
    W1 = W1.copy(); b1 = b1.copy(); W2 = W2.copy(); b2 = b2.copy(); w3 = w3.copy()
    W1[0] = [1.01, 2.0]
    b1[0] = 0.1
    W2[0, 0] = 0.51
    b2[0] = 0.25
    w3[0] = 2.0
    w3[5] = 9.0  # a dead unit: its input is zero
    w = mlp_ref.pack(W1, b1, W2, b2, w3, -1.0)
    x = np.array([[3.0, -1.0], [0.3, 0.0]])
    got = mlp_ref.forward_bf16(w, x, bands)
    assert got[0] == 3.305908203125 and got[1] == -0.5
    # the float64 network differs on sample 0: every rounding step matters
    assert mlp_ref.forward(w, x, bands)[0][0] != got[0]


def test_forward_f16_equals_forward_on_f16_exact_inputs():
    """Values that half precision holds exactly (multiples of 1/16 below 8):
    every rounding step is the identity, so the FP16 pass equals float64."""
    bands, n = 5, 40
    rng = np.random.default_rng(9)
    P = mlp_ref.param_count(bands)
    w = rng.integers(-4, 5, P) / 4.0
    W1, b1, W2, b2, w3, b3 = mlp_ref.unpack(w, bands)
    b1[:] = rng.integers(-8, 9, b1.shape) / 16.0
    w = mlp_ref.pack(W1, b1, W2, b2, w3, b3)
    x = rng.integers(-4, 5, (bands, n)) / 4.0
    z, h1, _ = mlp_ref.forward(w, x, bands)
    assert np.array_equal(mlp_ref.f16(h1), h1)
    assert np.abs(h1).max() > 0
    assert np.array_equal(mlp_ref.forward_f16(w, x, bands), z)


def test_forward_f16_hand_worked():
    """One live unit per layer. Half precision keeps 11 significant bits:
    f16(1.0001) = 1.0 (the next half above 1 is 1 + 2^-10), f16(0.3) =
    1229/4096 = 0.300048828125, f16(0.51) = 1044/2048 = 0.509765625.
    Sample 0, x = (3, 0.3): W1.x = 3 + 2*1229/4096 = 3.60009765625;
      + b1 = 3.70009765625 -> f16 1894/512 = 3.69921875; W2 h1 + b2 =
      (1044/2048)(1894/512) + 1/4 = 2.13573455810546875; logit = 2*that - 1
      = 214399/65536 = 3.2714691162109375.
    Sample 1, x = (-1, 0): relu 0; logit = 2*0.25 - 1 = -0.5."""
    bands = 2
    w = np.zeros(mlp_ref.param_count(bands))
    W1, b1, W2, b2, w3, b3 = mlp_ref.unpack(w, bands)
    W1 = W1.copy(); b1 = b1.copy(); W2 = W2.copy(); b2 = b2.copy(); w3 = w3.copy()
    W1[0] = [1.0001, 2.0]
    b1[0] = 0.1
    W2[0, 0] = 0.51
    b2[0] = 0.25
    w3[0] = 2.0
    w = mlp_ref.pack(W1, b1, W2, b2, w3, -1.0)
    x = np.array([[3.0, -1.0], [0.3, 0.0]])
    got = mlp_ref.forward_f16(w, x, bands)
    assert got[1] == -0.5
    assert got[0] == 214399 / 65536
    assert mlp_ref.forward(w, x, bands)[0][0] != got[0]
