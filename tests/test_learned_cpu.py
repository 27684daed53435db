"""CPU checks of the learned selector's oracle network (oracle/mlp_cpu.py).

The oracle is the checker of the HIP training kernels, so it is pinned here
by the mathematics: its analytic gradient against central finite
differences of its own loss in float64, planted mistakes that such a check
must catch, and the closed-form first Adam step.
"""
import numpy as np
import pytest

from oracle import mlp_cpu as M


def _batch(seed, B=64):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, M.N_IN))
    y = (rng.uniform(size=B) < 0.3).astype(float)
    theta = M.init_params(seed).astype(np.float64)
    theta[M.OFF_B1:M.OFF_W2] = rng.uniform(-0.1, 0.1, M.N_HID)
    theta[M.OFF_B2:M.OFF_W3] = rng.uniform(-0.1, 0.1, M.N_HID)
    theta[M.OFF_B3] = 0.05
    return theta, X, y


def _fd(theta, X, y, pw, idx, h=1e-6):
    g = np.empty(idx.size)
    for k, i in enumerate(idx):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        g[k] = (M.loss(tp, X, y, pw) - M.loss(tm, X, y, pw)) / (2 * h)
    return g


@pytest.mark.parametrize("pw", [1.0, 4.0])
def test_gradient_matches_finite_differences(pw):
    theta, X, y = _batch(0)
    L, g = M.loss_and_grad(theta, X, y, pw)
    assert abs(L - M.loss(theta, X, y, pw)) < 1e-12
    rng = np.random.default_rng(1)
    # every block of the layout, including the biases and the output bias
    idx = np.unique(np.concatenate([
        rng.integers(M.OFF_W1, M.OFF_B1, 12), rng.integers(M.OFF_B1, M.OFF_W2, 6),
        rng.integers(M.OFF_W2, M.OFF_B2, 16), rng.integers(M.OFF_B2, M.OFF_W3, 6),
        rng.integers(M.OFF_W3, M.OFF_B3, 6), [M.OFF_B3]]))
    fd = _fd(theta, X, y, pw, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-9)


def test_planted_mistakes_are_caught():
    """A dropped bias gradient or a transposed W2 must fail the FD check."""
    theta, X, y = _batch(2)
    _, g = M.loss_and_grad(theta, X, y, 2.0)
    idx_b = np.arange(M.OFF_B2, M.OFF_B2 + 8)
    fd_b = _fd(theta, X, y, 2.0, idx_b)
    bad = g.copy()
    bad[M.OFF_B2:M.OFF_W3] = 0.0
    assert not np.allclose(bad[idx_b], fd_b, rtol=1e-5, atol=1e-9)
    W2g = g[M.OFF_W2:M.OFF_B2].reshape(M.N_HID, M.N_HID)
    bad = g.copy()
    bad[M.OFF_W2:M.OFF_B2] = W2g.T.ravel()
    idx_w = np.arange(M.OFF_W2 + 1, M.OFF_W2 + 200, 7)
    fd_w = _fd(theta, X, y, 2.0, idx_w)
    assert np.allclose(g[idx_w], fd_w, rtol=1e-5, atol=1e-9)
    assert not np.allclose(bad[idx_w], fd_w, rtol=1e-5, atol=1e-9)


def test_adam_first_step_closed_form():
    """At step 1 the bias-corrected update is lr * g / (|g| + eps) -> +-lr."""
    rng = np.random.default_rng(3)
    theta = rng.standard_normal(50)
    grad = rng.choice([-1.0, 1.0], 50) * rng.uniform(0.5, 2.0, 50)  # eps/|g| <= 2e-8
    grad[5] = 0.0
    new, m1, m2 = M.adam_step(theta, grad, np.zeros(50), np.zeros(50), 1, lr=1e-2)
    step = theta - new
    nz = grad != 0
    assert np.allclose(np.abs(step[nz]), 1e-2, rtol=1e-6)
    assert np.all(np.sign(step[nz]) == np.sign(grad[nz]))
    assert step[5] == 0.0


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.0078125, 1.01171875, -3.14159265, 0.0], dtype=np.float32)
    r = M.bf16_round(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even mantissa (1.0);
    # 1 + 3*2^-8 ties to 1 + 2^-6 (even)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == 1.015625
    assert abs(r[4] - np.float32(-3.140625)) == 0.0 and r[5] == 0.0


def test_bf16_forward_close_to_fp32():
    theta, X, _ = _batch(4, B=256)
    z32, _ = M.forward(theta, X)
    zb = M.forward_bf16(theta.astype(np.float32), X.astype(np.float32))
    assert np.max(np.abs(zb - z32)) <= 3e-2 * (1.0 + np.max(np.abs(z32)))
