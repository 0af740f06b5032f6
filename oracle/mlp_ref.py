"""CPU oracle for the learned pre-screen (test infrastructure only).

The reference has no learned part, so this is not a restatement of any
reference file: it is the float64 numpy statement of the network the library
trains and runs on the GPU, used by the tests as the checker for the GPU's
forward pass, gradients and Adam step. It is pinned by central finite
differences of its own loss (tests/test_mlp_oracle.py), which a wrong sign,
a dropped term or a transposed operand in the backward pass would fail.

Network: x (B, centred) -> relu(W1 x + b1) (64) -> relu(W2 h1 + b2) (32) ->
w3 . h2 + b3 (logit). Loss per sample: pos_weight * y * softplus(-z) +
(1 - y) * softplus(z), averaged over the batch. Parameters packed as
W1[64][B] | b1[64] | W2[32][64] | b2[32] | w3[32] | b3.
"""
import numpy as np

H1, H2 = 64, 32


def param_count(bands):
    return H1 * bands + H1 + H2 * H1 + H2 + H2 + 1


def unpack(w, bands):
    w = np.asarray(w, np.float64)
    o = 0
    W1 = w[o:o + H1 * bands].reshape(H1, bands); o += H1 * bands
    b1 = w[o:o + H1]; o += H1
    W2 = w[o:o + H2 * H1].reshape(H2, H1); o += H2 * H1
    b2 = w[o:o + H2]; o += H2
    w3 = w[o:o + H2]; o += H2
    b3 = w[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [b3]])


def softplus(z):
    return np.logaddexp(0.0, z)


def forward(w, x, bands):
    """x: (B, n) float64 band-major samples -> (logit[n], h1, h2)."""
    W1, b1, W2, b2, w3, b3 = unpack(w, bands)
    h1 = np.maximum(W1 @ x + b1[:, None], 0.0)
    h2 = np.maximum(W2 @ h1 + b2[:, None], 0.0)
    return w3 @ h2 + b3, h1, h2


def loss_and_grad(w, x, y, pos_weight, bands):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = x.shape[1]
    W1, b1, W2, b2, w3, b3 = unpack(w, bands)
    z, h1, h2 = forward(w, x, bands)
    loss = np.mean(pos_weight * y * softplus(-z) + (1.0 - y) * softplus(z))
    p = 1.0 / (1.0 + np.exp(-z))
    dz = (pos_weight * y * (p - 1.0) + (1.0 - y) * p) / n
    gw3 = h2 @ dz
    gb3 = dz.sum()
    d2 = np.outer(w3, dz) * (h2 > 0)
    gW2 = d2 @ h1.T
    gb2 = d2.sum(axis=1)
    d1 = (W2.T @ d2) * (h1 > 0)
    gW1 = d1 @ x.T
    gb1 = d1.sum(axis=1)
    return loss, pack(gW1, gb1, gW2, gb2, gw3, gb3)


def adam_step(w, m, v, g, t, lr, beta1, beta2, eps):
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mhat = m / (1.0 - beta1 ** t)
    vhat = v / (1.0 - beta2 ** t)
    return w - lr * mhat / (np.sqrt(vhat) + eps), m, v


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def f16(a):
    """Round to IEEE half precision (nearest, ties to even), as float64."""
    return np.asarray(a, np.float64).astype(np.float16).astype(np.float64)


def forward_f16(w, x, bands):
    """The screen fused into the statistics pass: FP16 weights, inputs and
    layer-1 activations, FP32-class accumulation (here float64), FP32
    biases. `x` is the pixel minus the shift, already rounded to FP32."""
    W1, b1, W2, b2, w3, b3 = unpack(w, bands)
    h1 = np.maximum(f16(W1) @ f16(x) + b1[:, None], 0.0)
    h2 = np.maximum(f16(W2) @ f16(h1) + b2[:, None], 0.0)
    return w3 @ h2 + b3


def forward_bf16(w, x, bands):
    """The inference kernel's arithmetic: bf16 weights, inputs and layer-1
    activations, FP32-class accumulation (here float64), FP32 biases."""
    W1, b1, W2, b2, w3, b3 = unpack(w, bands)
    h1 = np.maximum(bf16(W1) @ bf16(x) + b1[:, None], 0.0)
    h2 = np.maximum(bf16(W2) @ bf16(h1) + b2[:, None], 0.0)
    return w3 @ h2 + b3
