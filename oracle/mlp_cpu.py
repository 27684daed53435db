"""ORACLE -- test infrastructure only.

CPU restatement of the learned column selector's network: a per-column MLP
8 -> 64 -> 64 -> 1 (ReLU hidden layers, one logit) with the feature
normalisation, the weighted binary cross-entropy loss, its gradient and the
Adam update that the HIP kernels implement.  The reference has no learned
component (its selection is the rule of psipm_stable_preconditioner.py:
293-371), so nothing here follows a reference file; the network is pinned by
the mathematics instead:

* ``loss_and_grad`` is checked against central finite differences of
  ``loss`` in float64 (tests/test_learned_cpu.py), with planted mistakes
  (a dropped bias gradient, a transposed weight) that the check catches;
* ``adam_step`` follows Kingma & Ba (2015), Algorithm 1, checked on a
  closed-form first step (|update| = lr for every nonzero gradient).

Only ``tests/`` may import this module, as the checker of the GPU kernels.

Parameter layout (float32, one flat vector of ``N_PARAMS``), shared with the
device code: W1 (8 x 64, row-major: W1[i, j] at i*64 + j), b1 (64),
W2 (64 x 64, row-major), b2 (64), w3 (64), b3 (1).
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 8, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAMS = OFF_B3 + 1


def unpack(theta):
    t = np.asarray(theta)
    return (t[OFF_W1:OFF_B1].reshape(N_IN, N_HID), t[OFF_B1:OFF_W2],
            t[OFF_W2:OFF_B2].reshape(N_HID, N_HID), t[OFF_B2:OFF_W3],
            t[OFF_W3:OFF_B3], t[OFF_B3])


def forward(theta, X, dtype=np.float64):
    """Logits of the rows of X (already normalised features, B x 8)."""
    W1, b1, W2, b2, w3, b3 = (np.asarray(a, dtype=dtype) for a in unpack(theta))
    X = np.asarray(X, dtype=dtype)
    h1 = np.maximum(X @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ w3 + b3, (h1, h2)


def loss(theta, X, y, pos_weight=1.0):
    """Mean weighted BCE-with-logits: -(pw y log s(z) + (1-y) log(1-s(z)))."""
    z, _ = forward(theta, X)
    y = np.asarray(y, dtype=np.float64)
    # log s(z) = -softplus(-z), log(1 - s(z)) = -softplus(z)
    sp_pos = np.logaddexp(0.0, -z)
    sp_neg = np.logaddexp(0.0, z)
    return float(np.mean(pos_weight * y * sp_pos + (1.0 - y) * sp_neg))


def loss_and_grad(theta, X, y, pos_weight=1.0):
    """Loss and its gradient with respect to the flat parameter vector."""
    W1, b1, W2, b2, w3, b3 = (np.asarray(a, dtype=np.float64) for a in unpack(theta))
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    B = X.shape[0]
    a1 = X @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3 + b3
    sig = 1.0 / (1.0 + np.exp(-z))
    L = float(np.mean(pos_weight * y * np.logaddexp(0.0, -z) + (1.0 - y) * np.logaddexp(0.0, z)))
    # d/dz of pw y softplus(-z) + (1-y) softplus(z) = (1-y) s - pw y (1-s)
    dz = ((1.0 - y) * sig - pos_weight * y * (1.0 - sig)) / B
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = h2.T @ dz
    g[OFF_B3] = dz.sum()
    d2 = np.outer(dz, w3) * (a2 > 0.0)
    g[OFF_W2:OFF_B2] = (h1.T @ d2).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (a1 > 0.0)
    g[OFF_W1:OFF_B1] = (X.T @ d1).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(axis=0)
    return L, g


def adam_step(theta, grad, m1, m2, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Alg. 1); returns new (theta, m1, m2)."""
    m1 = beta1 * m1 + (1.0 - beta1) * grad
    m2 = beta2 * m2 + (1.0 - beta2) * grad * grad
    mh = m1 / (1.0 - beta1 ** step)
    vh = m2 / (1.0 - beta2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m1, m2


def init_params(seed=0):
    """He-uniform initialisation, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    t = np.zeros(N_PARAMS, dtype=np.float32)
    lim1, lim2, lim3 = np.sqrt(6.0 / N_IN), np.sqrt(6.0 / N_HID), np.sqrt(3.0 / N_HID)
    t[OFF_W1:OFF_B1] = rng.uniform(-lim1, lim1, N_IN * N_HID)
    t[OFF_W2:OFF_B2] = rng.uniform(-lim2, lim2, N_HID * N_HID)
    t[OFF_W3:OFF_B3] = rng.uniform(-lim3, lim3, N_HID)
    return t


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest-even) and back."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def forward_bf16(theta, X):
    """The inference kernel's arithmetic: bf16 inputs, weights and hidden
    activations, float32 accumulation and biases."""
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(theta, dtype=np.float32))
    Xb = bf16_round(X).astype(np.float32)
    h1 = np.maximum(Xb @ bf16_round(W1) + b1, 0.0).astype(np.float32)
    h2 = np.maximum(bf16_round(h1) @ bf16_round(W2) + b2, 0.0).astype(np.float32)
    return (h2 @ w3 + b3).astype(np.float32)
