"""The reduced Newton system of the homogeneous self-dual (HSD) model on the
GPU KKT object -- the reference's second caller of the same preconditioner,
psipm_v3_hsd.py:894-1088 (``HSDNewtonSystem``).

For one Newton matrix H = A D A' + delta I the reduced direction needs three
solves with H: the step's own row right-hand side f and two that depend only
on the matrix, h = b + A D c_direction and the theta row (V3:1004-1019).
The reference runs three scalar PCGs and caches the last two for the
corrector directions.  Here the first direction solves f and h together in
one two-column block PCG (``solve_pair``: every read of A and of the factor
serves both) and the theta row alone; later directions of the same matrix
solve f alone and reuse the cached pair.  The 2 x 2 border system for
(dtau, dtheta), its near-singularity guard and the back-substitution for
dx, dz, dw, dkappa follow V3:1020-1088.

The HSD outer loop itself (state updates, step lengths, centring) stays with
the caller.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

EPS = float(np.finfo(np.float64).eps)


@dataclass
class HSDState:
    """The iterate fields the Newton system reads (V3:204-219)."""
    x: np.ndarray
    z: np.ndarray
    w: np.ndarray
    tau: float
    kappa: float


@dataclass
class HSDEmbedding:
    """Fixed residual data of the extended model (V3:222-229)."""
    bar_b: np.ndarray
    bar_c: np.ndarray
    bar_g: float


class HSDNewtonSystem:
    """Reduced HSD Newton system for one Newton matrix (V3:894-965).

    ``kkt`` is a GPU KKT object (``HipSpectralStableColumnPreconditioner`` or
    ``HipTreePreconditioner``); it is factorised here with theta = 1 / D and
    its tolerance fields set to ``pcg_tolerance``.  ``pair=False`` solves the
    three rows one by one, as the reference does.
    """

    def __init__(self, kkt, A, c, upper, b, D, delta, eta, state: HSDState, embedding: HSDEmbedding,
                 pcg_tolerance=1e-8, zeta=1e-8, pair=True):
        self.kkt, self.A = kkt, A.tocsc()
        self.c = np.asarray(c, dtype=float)
        self.upper = np.asarray(upper, dtype=float)
        self.b = np.asarray(b, dtype=float)
        self.D = np.asarray(D, dtype=float)
        self.delta, self.eta = float(delta), float(eta)
        self.state, self.embedding = state, embedding
        self.pair = bool(pair)
        self.m, self.n = self.A.shape
        kkt.cg_tol = kkt.cg_tol_max = kkt.current_cg_tol = float(pcg_tolerance)
        kkt.factorize(1.0 / np.maximum(self.D, np.finfo(float).tiny), self.delta)
        boxed = np.isfinite(self.upper)
        self.boxed = boxed
        ub = np.where(boxed, self.upper, 0.0)
        self.slack = np.ones_like(self.c)
        self.slack[boxed] = self.upper[boxed] * state.tau - state.x[boxed]
        ratio = np.zeros_like(self.c)
        ratio[boxed] = state.w[boxed] / self.slack[boxed]
        self.c_direction = self.c - ratio * ub
        self.c_gap = self.c + ratio * ub
        Dc = self.D * embedding.bar_c
        self.h = self.b + self._A(self.D * self.c_direction)
        self.g = self.b - self._A(self.D * self.c_gap)
        ADc = self._A(Dc)
        self.theta_row = embedding.bar_b + ADc
        self.normal_row = ADc - embedding.bar_b
        self.alpha = (float(self.c_gap @ (self.D * self.c_direction)) + float(np.sum(ratio[boxed] * ub[boxed] ** 2))
                      + state.kappa / state.tau + self.eta)
        self.beta = embedding.bar_g - float(self.c_gap @ Dc)
        self.gamma = -float(embedding.bar_c @ (self.D * self.c_direction)) - embedding.bar_g
        self.lambda_theta = float(embedding.bar_c @ Dc) + float(zeta)
        self._h_solution = None
        self._theta_solution = None
        self.solves = 0
        self.pair_solves = 0

    def _A(self, v):
        mv = getattr(self.kkt, "matvec", None)
        return mv(v) if mv is not None else self.A @ v

    def _AT(self, y):
        rmv = getattr(self.kkt, "rmatvec", None)
        return rmv(y) if rmv is not None else self.A.T @ y

    def _solve_one(self, rhs):
        self.solves += 1
        _, dy = self.kkt.solve(np.zeros(self.n), rhs)
        return dy

    def direction(self, rhs_d, rhs_p, rhs_g, rhs_n, rhs_c, rhs_upper, rhs_t):
        """(dx, dy, dz, dw, dtau, dtheta, dkappa, ok) of V3:985-1088."""
        st, emb, boxed = self.state, self.embedding, self.boxed
        up = np.zeros_like(rhs_d)
        up[boxed] = rhs_upper[boxed] / self.slack[boxed]
        q = rhs_d + rhs_c / st.x - up
        Dq = self.D * q
        f = -rhs_p - self._A(Dq)
        ok = True
        try:
            if self._h_solution is None and self.pair:
                (_, u), (_, hs) = self.kkt.solve_pair(np.zeros(self.n), f, np.zeros(self.n), self.h)
                self.solves += 2
                self.pair_solves += 1
                self._h_solution = hs
            else:
                u = self._solve_one(f)
                if self._h_solution is None:
                    self._h_solution = self._solve_one(self.h)
            if self._theta_solution is None:
                self._theta_solution = self._solve_one(self.theta_row)
        except Exception as exc:  # the KKT object's failure ladder was exhausted
            if not isinstance(exc, getattr(self.kkt, "error_class", Exception)):
                raise
            return None, None, None, None, 0.0, 0.0, 0.0, False
        hs, ts = self._h_solution, self._theta_solution
        gap_rhs = (rhs_g + float(self.c_gap @ Dq) + float(self.upper[boxed] @ (rhs_upper[boxed] / self.slack[boxed]))
                   + rhs_t / st.tau - float(self.g @ u))
        normal_rhs = rhs_n - float(emb.bar_c @ Dq) - float(self.normal_row @ u)
        border = np.array([[self.alpha + float(self.g @ hs), self.beta - float(self.g @ ts)],
                           [self.gamma + float(self.normal_row @ hs), self.lambda_theta - float(self.normal_row @ ts)]])
        border_rhs = np.array([gap_rhs, normal_rhs])
        scale = max(float(np.linalg.norm(border, ord=np.inf)), 1.0)
        if (not np.all(np.isfinite(border)) or not np.all(np.isfinite(border_rhs))
                or abs(float(np.linalg.det(border))) <= 64.0 * EPS * scale * scale):
            ok = False
            border = border + 64.0 * EPS * scale * np.eye(2)
        try:
            dtau, dtheta = np.linalg.solve(border, border_rhs)
        except np.linalg.LinAlgError:
            dtau, dtheta = np.linalg.lstsq(border, border_rhs, rcond=None)[0]
            ok = False
        dy = u + hs * dtau - ts * dtheta
        dx = self.D * (q + self._AT(dy) - self.c_direction * dtau + emb.bar_c * dtheta)
        dz = (rhs_c - st.z * dx) / st.x
        dw = np.zeros_like(dx)
        ds = np.zeros_like(dx)
        ds[boxed] = self.upper[boxed] * dtau - dx[boxed]
        dw[boxed] = (rhs_upper[boxed] - st.w[boxed] * ds[boxed]) / self.slack[boxed]
        dkappa = (rhs_t - st.kappa * dtau) / st.tau
        finite = all(np.all(np.isfinite(v)) for v in (dx, dy, dz, dw, np.asarray([dtau, dtheta, dkappa])))
        return dx, dy, dz, dw, float(dtau), float(dtheta), float(dkappa), bool(ok and finite)
