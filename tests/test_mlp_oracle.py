"""Pins the FP64 schedule-network oracle (oracle/mlp.py) on the CPU.

The network has no reference counterpart, so the oracle is pinned by
mathematics: the analytic backward must match FP64 central differences in
every layer, the forward must reduce to hand-computable cases, and one Adam
step must give its closed form.  A planted mistake (a dropped ReLU mask in
the backward) must fail the finite-difference check.
"""
import numpy as np

from oracle import mlp as M


def _params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return M.fp16(rng.standard_normal(M.NPARAMS).astype(np.float32) * scale)


def _layer_indices(rng, per_layer=12):
    idx, o = [], 0
    for s in M.SHAPES:
        k = int(np.prod(s))
        idx.extend((o + rng.choice(k, size=min(per_layer, k), replace=False)).tolist())
        o += k
    return np.array(idx)


def test_parameter_count_matches_library_header():
    assert M.NPARAMS == 5508


def test_backward_matches_central_differences_in_every_layer():
    rng = np.random.default_rng(1)
    p = _params(1)
    x = rng.uniform(0, 1, size=(37, M.IN))
    y = rng.integers(0, M.OUT, size=37)
    idx = _layer_indices(rng)
    g = M.grad(p, x, y)[idx]
    fd = M.grad_fd(p, x, y, idx)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-9)


def test_forward_special_cases():
    x = np.zeros((1, M.IN))
    x[0, 3] = 1.0
    # zero weights: logits are the output biases
    p = np.zeros(M.NPARAMS)
    W1, b1, W2, b2, W3, b3 = M.unpack(np.arange(M.NPARAMS, dtype=np.float64))
    o_b3 = int(b3[0])
    p[o_b3:o_b3 + 4] = [0.5, -1.0, 2.0, 0.25]
    z, _ = M.forward(p, x)
    np.testing.assert_array_equal(z[0], [0.5, -1.0, 2.0, 0.25])
    # a single path x3 -> h1[5] -> h2[7] -> logit 2: product of the weights
    q = np.zeros(M.NPARAMS)
    q[int(W1[5, 3])] = 0.5
    q[int(W2[7, 5])] = 3.0
    q[int(W3[2, 7])] = -2.0
    z, _ = M.forward(q, x)
    np.testing.assert_array_equal(z[0], [0.0, 0.0, -3.0, 0.0])
    # ReLU blocks a negative pre-activation
    q[int(W1[5, 3])] = -0.5
    z, _ = M.forward(q, x)
    np.testing.assert_array_equal(z[0], [0.0, 0.0, 0.0, 0.0])


def test_loss_of_uniform_logits_is_log4():
    p = np.zeros(M.NPARAMS)
    x = np.random.default_rng(2).uniform(size=(5, M.IN))
    assert abs(M.loss(p, x, np.array([0, 1, 2, 3, 0])) - np.log(4.0)) < 1e-15


def test_fp16_rounding_is_round_to_nearest_even():
    v = np.array([1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 65504.0, 1e-8], np.float32)
    np.testing.assert_array_equal(M.fp16(v), [1.0, 1.0 + 2.0 ** -9, 65504.0, 0.0])


def test_adam_first_step_closed_form():
    # step 1: m/(1-b1) = g and v/(1-b2) = g^2, so the update is -lr*g/(|g|+eps)
    g = np.array([0.5, -2.0, 1e-3])
    w, m1, m2 = M.adam_step(np.zeros(3), np.zeros(3), np.zeros(3), g, 1, 0.01)
    np.testing.assert_allclose(w, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_planted_missing_relu_mask_is_caught():
    rng = np.random.default_rng(3)
    p = _params(3)
    x = rng.uniform(0, 1, size=(16, M.IN))
    y = rng.integers(0, M.OUT, size=16)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    z3, (xx, z1, h1, z2, h2) = M.forward(p, x)
    e = np.exp(z3 - z3.max(axis=1, keepdims=True))
    dz3 = e / e.sum(axis=1, keepdims=True)
    dz3[np.arange(16), y] -= 1.0
    dz3 /= 16
    dz2_wrong = dz3 @ W3  # mask dropped
    gW2_wrong = (dz2_wrong.T @ h1).ravel()
    o = M.H * M.IN + M.H
    idx = o + rng.choice(M.H * M.H, size=40, replace=False)
    fd = M.grad_fd(p, x, y, idx)
    assert not np.allclose(gW2_wrong[idx - o], fd, rtol=1e-3, atol=1e-7)
    np.testing.assert_allclose(M.grad(p, x, y)[idx], fd, rtol=1e-5, atol=1e-9)
