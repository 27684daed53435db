"""ORACLE -- test infrastructure only.

CPU FP64 restatement of the reference's second caller of the normal-
equations preconditioner: the reduced Newton system of the homogeneous
self-dual (HSD) model, psipm_v3_hsd.py:894-1088 (``HSDNewtonSystem``).
Only ``tests/`` may import it, as the checker of the GPU Newton system.

* ``newton_data``  -- V3:932-965: upper-bound slacks and ratios, the cost
  vectors c_direction / c_gap, h = b + A D c_direction, g = b - A D c_gap,
  the theta and normal rows, and the border scalars alpha, beta, gamma,
  lambda_theta.
* ``direction``    -- V3:985-1088 with exact solves of H = A D A' + delta I
  (scipy's sparse LU): u = H^-1 f, H^-1 h, H^-1 theta_row, the 2 x 2 border
  system for (dtau, dtheta), then dy, dx, dz, dw, dkappa.
* ``bordered_solve`` -- the pin: the same (dy, dtau, dtheta) from the dense
  (m + 2) x (m + 2) bordered system
      [ H        -h      theta_row ] [dy    ]   [ f ]
      [ g'        alpha   beta     ] [dtau  ] = [ G ]
      [ normal'   gamma   lambda   ] [dtheta]   [ N ]
  that the reduced elimination solves (G and N are the border right-hand
  sides before the eliminated u terms), with numpy's dense solver.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def newton_data(A, c, upper, b, D, delta, eta, state, emb, zeta):
    boxed = np.isfinite(upper)
    upper0 = np.where(boxed, upper, 0.0)
    slack = np.ones_like(c)
    slack[boxed] = upper[boxed] * state.tau - state.x[boxed]
    ratio = np.zeros_like(c)
    ratio[boxed] = state.w[boxed] / slack[boxed]
    c_dir = c - ratio * upper0
    c_gap = c + ratio * upper0
    d = SimpleNamespace(boxed=boxed, slack=slack, ratio=ratio, c_dir=c_dir, c_gap=c_gap)
    d.h = b + A @ (D * c_dir)
    d.g = b - A @ (D * c_gap)
    d.theta_row = emb.bar_b + A @ (D * emb.bar_c)
    d.normal_row = A @ (D * emb.bar_c) - emb.bar_b
    d.alpha = float(c_gap @ (D * c_dir)) + float(np.sum(ratio[boxed] * upper[boxed] ** 2)) + state.kappa / state.tau + eta
    d.beta = emb.bar_g - float(c_gap @ (D * emb.bar_c))
    d.gamma = -float(emb.bar_c @ (D * c_dir)) - emb.bar_g
    d.lam = float(emb.bar_c @ (D * emb.bar_c)) + zeta
    d.H = (A @ sp.diags(D) @ A.T + delta * sp.identity(A.shape[0])).tocsc()
    return d


def _reduced_rhs(A, D, upper, state, d, rhs):
    rhs_d, rhs_p, rhs_g, rhs_n, rhs_c, rhs_upper, rhs_t = rhs
    up = np.zeros_like(rhs_d)
    up[d.boxed] = rhs_upper[d.boxed] / d.slack[d.boxed]
    q = rhs_d + rhs_c / state.x - up
    f = -rhs_p - A @ (D * q)
    G = (rhs_g + float(d.c_gap @ (D * q)) + float(upper[d.boxed] @ (rhs_upper[d.boxed] / d.slack[d.boxed]))
         + rhs_t / state.tau)
    return q, f, G


def direction(A, c, upper, b, D, delta, eta, state, emb, zeta, rhs):
    """(dx, dy, dz, dw, dtau, dtheta, dkappa) of V3:985-1088, exact solves."""
    d = newton_data(A, c, upper, b, D, delta, eta, state, emb, zeta)
    rhs_d, rhs_p, rhs_g, rhs_n, rhs_c, rhs_upper, rhs_t = rhs
    q, f, G = _reduced_rhs(A, D, upper, state, d, rhs)
    lu = spla.splu(d.H)
    u, hs, ts = lu.solve(f), lu.solve(d.h), lu.solve(d.theta_row)
    N = rhs_n - float(emb.bar_c @ (D * q))
    border = np.array([[d.alpha + d.g @ hs, d.beta - d.g @ ts],
                       [d.gamma + d.normal_row @ hs, d.lam - d.normal_row @ ts]])
    dtau, dtheta = np.linalg.solve(border, np.array([G - d.g @ u, N - d.normal_row @ u]))
    dy = u + hs * dtau - ts * dtheta
    dx = D * (q + A.T @ dy - d.c_dir * dtau + emb.bar_c * dtheta)
    dz = (rhs_c - state.z * dx) / state.x
    dw = np.zeros_like(dx)
    ds = np.zeros_like(dx)
    ds[d.boxed] = upper[d.boxed] * dtau - dx[d.boxed]
    dw[d.boxed] = (rhs_upper[d.boxed] - state.w[d.boxed] * ds[d.boxed]) / d.slack[d.boxed]
    dkappa = (rhs_t - state.kappa * dtau) / state.tau
    return dx, dy, dz, dw, float(dtau), float(dtheta), float(dkappa)


def bordered_solve(A, c, upper, b, D, delta, eta, state, emb, zeta, rhs):
    """(dy, dtau, dtheta) from the dense bordered system (small cases)."""
    d = newton_data(A, c, upper, b, D, delta, eta, state, emb, zeta)
    q, f, G = _reduced_rhs(A, D, upper, state, d, rhs)
    N = rhs[3] - float(emb.bar_c @ (D * q))
    m = A.shape[0]
    K = np.zeros((m + 2, m + 2))
    K[:m, :m] = d.H.toarray()
    K[:m, m] = -d.h
    K[:m, m + 1] = d.theta_row
    K[m, :m] = d.g
    K[m, m], K[m, m + 1] = d.alpha, d.beta
    K[m + 1, :m] = d.normal_row
    K[m + 1, m], K[m + 1, m + 1] = d.gamma, d.lam
    sol = np.linalg.solve(K, np.r_[f, G, N])
    return sol[:m], float(sol[m]), float(sol[m + 1])
