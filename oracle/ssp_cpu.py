"""ORACLE -- test infrastructure only.

CPU FP64 restatement of the reference's hot path: the spectral stable
active-column preconditioner and its PCG solve of the proximal normal
equations.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker or the
timed CPU baseline -- never as a product code path.

Restated (numpy/scipy, no import of the reference) from:

* psipm_active_pool.py:691-717   ``low_rank_columns``       (Woodbury column pick)
* psipm_active_pool.py:720-996   ``NormalEquations``         (Jacobi scaling,
  shift ladder, Woodbury core, ``_apply``)
* psipm_active_pool.py:999-1006  ``column_weighted_normal``
* psipm_stable_preconditioner.py:97-110  ``top_k``
* psipm_stable_preconditioner.py:113-438 stable core (candidates, backbone,
  scores, refresh, assemble, full matvec)
* psipm_spectral_preconditioner.py:119-140 ``ritz_estimate``
* psipm_spectral_preconditioner.py:143-752 spectral object (EW forcing,
  factorize decisions, lag bound, PCG with residual replacement, MINRES
  fallback, retry ladder)

The numeric factor is a library routine standing in for CHOLMOD (absent
here): LAPACK dense Cholesky (``scipy.linalg.cholesky``) when m is small,
SuperLU in symmetric mode with diagonal pivoting otherwise.  Any exact
factorisation of the same matrix gives the same P^-1 r up to rounding, which
is the quantity the parity tests compare.

Parity pinning: the preconditioner apply is pinned by the dense LAPACK
solve; the PCG is pinned by exact convergence on small SPD systems and by
scipy ``cg`` iteration agreement; the end-to-end answer by scipy's HiGHS.
``factor values`` (CHOLMOD's L itself) are parity unpinned: the reference
has no test fixing them and only P^-1 r enters the algorithm.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.linalg as la
import scipy.sparse as sp
import scipy.sparse.csgraph as csgraph
import scipy.sparse.linalg as spla


class KKTSolveError(RuntimeError):
    pass


SHIFTS = (0.0, 1e-14, 1e-12, 1e-10, 1e-8, 1e-6, 1e-4, 1e-2)


def low_rank_columns(A: sp.csc_matrix, max_columns=8, min_nnz=1024, clique_ratio=2.0):
    """Columns whose clique q(q-1)/2 dominates the matrix (PS:691-717)."""
    if max_columns <= 0 or A.shape[1] == 0:
        return np.empty(0, dtype=np.int64)
    q = np.diff(A.indptr).astype(float)
    edges = 0.5 * q * np.maximum(q - 1.0, 0.0)
    scale = max(float(A.nnz), float(A.shape[0]), 1.0)
    cand = np.flatnonzero((q >= max(1, int(min_nnz))) & (edges >= float(clique_ratio) * scale))
    if cand.size > max_columns:
        cand = cand[np.argpartition(edges[cand], -max_columns)[-max_columns:]]
    return np.sort(cand.astype(np.int64))


def column_weighted_normal(A: sp.csc_matrix, weights: np.ndarray) -> sp.csc_matrix:
    """A diag(weights) A^T (PS:999-1006)."""
    cols = np.repeat(np.arange(A.shape[1]), np.diff(A.indptr))
    Ad = sp.csc_matrix((A.data * np.sqrt(weights)[cols], A.indices, A.indptr), shape=A.shape)
    return (Ad @ Ad.T).tocsc()


def top_k(score: np.ndarray, count: int) -> np.ndarray:
    """Largest finite scores, ordered by (score desc, index asc) (SAP:97-110)."""
    count = min(max(0, int(count)), score.size)
    if count == 0:
        return np.empty(0, dtype=np.int64)
    finite = np.flatnonzero(np.isfinite(score))
    if finite.size <= count:
        chosen = finite
    else:
        chosen = finite[np.argpartition(score[finite], -count)[-count:]]
    return chosen[np.lexsort((chosen, -score[chosen]))].astype(np.int64)


def ritz_estimate(alphas, betas):
    """Extreme Ritz values of the CG Lanczos tridiagonal (SSP:119-140)."""
    k = len(alphas)
    if k == 0:
        return 1.0, 1.0, 1.0
    a = np.asarray(alphas)
    diag = np.empty(k)
    diag[0] = 1.0 / a[0]
    if k > 1:
        b = np.asarray(betas[: k - 1])
        diag[1:] = 1.0 / a[1:] + b / a[:-1]
        off = np.sqrt(np.maximum(b, 0.0)) / a[:-1]
        ev = la.eigvalsh_tridiagonal(diag, off, check_finite=False)
    else:
        ev = diag
    lo, hi = float(ev[0]), float(ev[-1])
    if not np.isfinite(lo) or lo <= 0.0:
        return lo, hi, np.inf
    return lo, hi, hi / lo


class _LibraryFactor:
    """Exact factor of an SPD matrix by a library routine (checker for CHOLMOD)."""

    dense_limit = 6000

    def __init__(self, M: sp.csc_matrix):
        m = M.shape[0]
        if m <= self.dense_limit:
            self.kind = "dense"
            self.L = la.cholesky(M.toarray(), lower=True, check_finite=True)
            self.nnz = int(np.count_nonzero(self.L))
        else:
            self.kind = "superlu"
            lu = spla.splu(
                M.tocsc(), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                options=dict(SymmetricMode=True),
            )
            d = lu.U.diagonal()
            if not np.all(np.isfinite(d)) or np.any(d <= 0.0):
                raise np.linalg.LinAlgError("not positive definite")
            self.lu = lu
            self.nnz = int(lu.L.nnz)

    def solve(self, r: np.ndarray) -> np.ndarray:
        if self.kind == "dense":
            return la.cho_solve((self.L, True), r, check_finite=False)
        return self.lu.solve(r)


@dataclass
class KrylovResult:
    solution: np.ndarray
    info: int
    iterations: int
    relative_residual: float
    condition_estimate: float
    spectral_error_estimate: float
    residual: np.ndarray
    replacements: int
    residual_checks: int
    stagnated: bool


class SpectralStableColumnPreconditionerCPU:
    """CPU oracle of the spectral stable active-column KKT object."""

    def __init__(
        self, A: sp.csc_matrix, protected: Optional[np.ndarray] = None,
        core_factor=1.0, max_core_factor=3.0, retention=0.9, factor_max_age=2,
        core_max_age=12, refresh_cg=30, fill_weight=0.2, cg_tol=1e-7, cg_tol_max=1e-2,
        forcing_factor=0.1, cg_maxit=200, retry_growth=1.35,
        full_preconditioner_row_threshold=1500, full_preconditioner_clique_limit=2e8,
        full_preconditioner_nnz_limit=10_000_000, structural_backbone=True,
        residual_score_weight=0.0, swap_patience=2, condition_limit=1e12,
        lag_spectral_ratio=1e8, residual_replacement=100, residual_gap_tolerance=0.1,
        stagnation_window=3, stagnation_reduction=0.8, corrector_tightening=1.0,
        minres_fallback=True, minres_maxit=400, minres_max_dimension=200_000,
        fail_tol=1e-6, low_rank_max_columns=8, low_rank_min_nnz=1024,
        low_rank_clique_ratio=2.0, **_ignored,
    ):
        if not 0.0 < cg_tol <= cg_tol_max < 1.0:
            raise ValueError("require 0 < cg_tol <= cg_tol_max < 1")
        self.A = A.tocsc()
        self.A.sum_duplicates()
        self.A.sort_indices()
        self.At = self.A.T.tocsr()
        self.m, self.n = self.A.shape
        self.fail_tol = float(fail_tol)
        self.low_rank_columns = low_rank_columns(
            self.A, low_rank_max_columns, low_rank_min_nnz, low_rank_clique_ratio)
        mask = np.ones(self.n, dtype=bool)
        mask[self.low_rank_columns] = False
        self.factor_columns = np.flatnonzero(mask)
        self.A_factor = self.A[:, self.factor_columns].tocsc()
        self.A_low_rank = self.A[:, self.low_rank_columns].tocsc()
        counts = np.diff(self.A_factor.indptr).astype(np.int64)
        coupling = np.flatnonzero(counts >= 2)
        self.candidates = self.factor_columns[coupling]
        self.A_coupling = self.A_factor[:, coupling].tocsc()
        sq = lambda M: sp.csc_matrix((np.square(M.data), M.indices, M.indptr), shape=M.shape)
        self.A2_coupling = sq(self.A_coupling)
        self.A2_factor = sq(self.A_factor)
        self.A2_low_rank = sq(self.A_low_rank)
        q = counts[coupling].astype(float)
        self.clique_burden = 0.5 * q * np.maximum(q - 1.0, 0.0)
        self.use_full_preconditioner = bool(
            self.m <= max(0, int(full_preconditioner_row_threshold))
            and self.A_factor.nnz <= max(0, int(full_preconditioner_nnz_limit))
            and float(self.clique_burden.sum()) <= float(full_preconditioner_clique_limit)
        )
        pmask = np.zeros(self.n, dtype=bool)
        if protected is not None:
            pmask[:] = np.asarray(protected, dtype=bool)
        self.protected_mandatory = self.candidates[pmask[self.candidates]]
        self.mandatory = self.protected_mandatory.copy()
        self.matching_columns = np.empty(0, dtype=np.int64)
        self.structural_backbone = bool(structural_backbone)
        requested = max(1, int(np.ceil(float(core_factor) * self.m)))
        maximum = max(requested, int(np.ceil(float(max_core_factor) * self.m)))
        if self.use_full_preconditioner:
            requested = maximum = self.candidates.size
        self.core_target = min(self.candidates.size, max(requested, self.mandatory.size))
        self.max_core_target = min(self.candidates.size, max(maximum, self.mandatory.size))
        self.retention = float(retention)
        self.factor_max_age = 0 if self.use_full_preconditioner else max(0, int(factor_max_age))
        self.core_max_age = max(1, int(core_max_age))
        self.refresh_cg = max(1, int(refresh_cg))
        self.fill_weight = float(fill_weight)
        self.cg_tol, self.cg_tol_max = float(cg_tol), float(cg_tol_max)
        self.forcing_factor = float(forcing_factor)
        self.current_cg_tol = self.cg_tol
        self.cg_maxit = max(1, int(cg_maxit))
        self.retry_growth = float(retry_growth)
        self.residual_score_weight = max(0.0, float(residual_score_weight))
        self.swap_patience = max(1, int(swap_patience))
        self.condition_limit = max(1.0, float(condition_limit))
        self.lag_spectral_ratio = max(1.0, float(lag_spectral_ratio))
        self.residual_replacement = max(0, int(residual_replacement))
        self.residual_gap_tolerance = max(np.sqrt(np.finfo(float).eps), float(residual_gap_tolerance))
        self.stagnation_window = max(2, int(stagnation_window))
        self.stagnation_reduction = float(stagnation_reduction)
        self.corrector_tightening = min(1.0, max(0.01, float(corrector_tightening)))
        self.minres_fallback = bool(minres_fallback)
        self.minres_maxit = max(1, int(minres_maxit))
        self.minres_max_dimension = max(0, int(minres_max_dimension))

        self.core = np.empty(0, dtype=np.int64)
        self._weights = None
        self._diag_drop = np.zeros(self.m)
        self.factor_age = self.factor_max_age
        self.core_age = self.core_max_age
        self._initialized = False
        self._last_cg = self.refresh_cg
        self._factor = None
        self._pattern = None
        self._dsc = None
        self._shift = 0.0
        self._lr_basis = None
        self._lr_chol = None
        self.theta = None
        self.delta = 0.0
        self._enrichment = None
        self._reference_weights = None
        self._reference_delta = None
        self._poor_streak = 0
        self._last_condition = 1.0
        self._last_spectral_error = 0.0
        self._last_stagnated = False
        self._rhs_index = 0
        self._nonlinear = None
        self._ew_tolerance = self.cg_tol_max
        # counters (names read by the driver)
        self.n_fact = self.n_solve = 0
        self.t_fact = self.t_solve = 0.0
        self.n_shifted = self.n_refine = self.n_symbolic = 0
        self.stats = {}
        self.core_sizes, self.cg_iterations, self.cg_tolerances = [], [], []
        self.cg_failures = self.pcg_retries = self.full_fallbacks = self.refreshes = 0
        self.retained_fractions, self.preconditioner_nnz = [], []
        self.condition_estimates, self.spectral_error_estimates, self.lag_spectral_bounds = [], [], []
        self.residual_replacements = self.residual_checks = self.stagnation_events = 0
        self.core_swaps = self.minres_fallbacks = self.minres_iterations = 0
        self.krylov_breakdowns = self.krylov_iteration_limits = 0

    # ---------------- IPM hooks (SSP:230-265) ----------------
    def set_ipm_context(self, res_d, res_p, res_n, tol):
        tiny = np.finfo(float).tiny
        current = max(float(res_n), float(np.linalg.norm(res_d)), float(np.linalg.norm(res_p)), tiny)
        prev = self._nonlinear
        self._nonlinear = current
        est = self.cg_tol_max if prev is None or not np.isfinite(prev) else 0.9 * (current / max(prev, tiny)) ** 1.5
        self._ew_tolerance = min(self.cg_tol_max, max(self.cg_tol, float(tol) * 0.1, est))
        self._rhs_index = 0

    def observe(self, x, z, w, s, mu):
        norm_mu = max(float(mu), 0.0) / (1.0 + max(float(mu), 0.0))
        self.current_cg_tol = min(self.cg_tol_max, max(self.cg_tol, min(self.forcing_factor * norm_mu, self._ew_tolerance)))

    def _linear_tolerance(self):
        tol = self.current_cg_tol
        if self._rhs_index:
            tol = max(self.cg_tol, self.corrector_tightening * tol)
        self._rhs_index += 1
        return float(tol)

    # ---------------- selection (SAP:293-371, SSP:307-394) ----------------
    def _structural_matching(self):
        if not self.structural_backbone or self.use_full_preconditioner or not self.m or not self.factor_columns.size:
            return np.empty(0, dtype=np.int64)
        g = self.A_factor.tocsr(copy=True)
        g.data.fill(1.0)
        local = csgraph.maximum_bipartite_matching(g, perm_type="column")
        glob = self.factor_columns[local[local >= 0]]
        pos = np.searchsorted(self.candidates, glob)
        ok = pos < self.candidates.size
        idx = np.flatnonzero(ok)
        ok[idx] &= self.candidates[pos[idx]] == glob[idx]
        return np.unique(glob[ok])

    def _update_backbone(self):
        if self.matching_columns.size:
            return
        match = self._structural_matching()
        if match.size:
            self.matching_columns = match
            self.mandatory = np.union1d(self.protected_mandatory, match)
            self.core_target = min(self.candidates.size, max(self.core_target, self.mandatory.size))
            self.max_core_target = min(self.candidates.size, max(self.max_core_target, self.mandatory.size))

    def _scores(self, weights, delta):
        if not self.candidates.size:
            return np.empty(0)
        diag = np.asarray(self.A2_factor @ weights[self.factor_columns]).ravel()
        if self.low_rank_columns.size:
            diag += np.asarray(self.A2_low_rank @ weights[self.low_rank_columns]).ravel()
        diag += max(float(delta), np.finfo(float).tiny)
        inv = 1.0 / np.maximum(diag, np.finfo(float).tiny)
        lev = weights[self.candidates] * np.asarray(self.A2_coupling.T @ inv).ravel()
        scores = lev / (1.0 + self.fill_weight * self.clique_burden)
        v = self._enrichment
        if self.residual_score_weight > 0.0 and scores.size and v is not None and v.size == self.m:
            nv = float(np.linalg.norm(v))
            if np.isfinite(nv) and nv > 0.0:
                v = v / nv
                prod = np.asarray(self.A_coupling.T @ v).ravel()
                dg = np.asarray(self.A2_coupling.T @ np.square(v)).ravel()
                impact = weights[self.candidates] * np.abs(np.square(prod) - dg)
                impact /= 1.0 + self.fill_weight * self.clique_burden
                s_sc = float(np.max(scores, initial=0.0))
                i_sc = float(np.max(impact, initial=0.0))
                if i_sc > 0.0 and np.isfinite(i_sc):
                    scores = scores + self.residual_score_weight * max(s_sc, 1.0) * (impact / i_sc)
        return scores

    def _refresh_core(self, weights, delta, stable):
        old = self.core.copy()
        self._update_backbone()
        if self.use_full_preconditioner:
            self.core = self.candidates.copy()
            self.retained_fractions.append(1.0)
        else:
            scores = self._scores(weights, delta)
            target = min(self.core_target, self.candidates.size)
            picked = []
            mand_local = np.searchsorted(self.candidates, self.mandatory)
            if self.mandatory.size:
                picked.append(mand_local)
            if stable and old.size and target:
                old_local = np.searchsorted(self.candidates, old)
                keep = min(old.size, max(0, int(np.floor(self.retention * target))))
                if keep:
                    picked.append(old_local[top_k(scores[old_local], keep)])
            fixed = np.unique(np.concatenate(picked)) if picked else np.empty(0, dtype=np.int64)
            if fixed.size > target:
                if self.mandatory.size >= target:
                    fixed = mand_local[:target]
                else:
                    optional = fixed[~np.isin(fixed, mand_local, assume_unique=True)]
                    fixed = np.concatenate([mand_local, optional[top_k(scores[optional], target - self.mandatory.size)]])
            remaining = target - fixed.size
            if remaining > 0:
                avail = scores.copy()
                avail[fixed] = -np.inf
                fixed = np.concatenate([fixed, top_k(avail, remaining)])
            new = np.sort(self.candidates[np.unique(fixed)])
            self.retained_fractions.append(
                float(np.intersect1d(old, new, assume_unique=True).size) / max(1, old.size) if old.size else 1.0)
            self.core = new
        self.refreshes += 1
        self.core_age = 0
        self._initialized = True
        if old.size and self.core.size:
            self.core_swaps += max(0, int(old.size - np.intersect1d(old, self.core, assume_unique=True).size))

    # ---------------- preconditioner matrix and factor ----------------
    def assemble(self, weights, delta):
        """P = A_C W_C A_C' + diag(max(d_full - d_core, 0)) + delta I (SAP:373-395)."""
        if self.core.size:
            matrix = column_weighted_normal(self.A[:, self.core].tocsc(), weights[self.core])
            loc = np.searchsorted(self.candidates, self.core)
            diag_core = np.asarray(self.A2_coupling[:, loc] @ weights[self.core]).ravel()
        else:
            matrix = sp.csc_matrix((self.m, self.m))
            diag_core = np.zeros(self.m)
        diag_full = np.asarray(self.A2_factor @ weights[self.factor_columns]).ravel()
        self._diag_drop = np.maximum(diag_full - diag_core, 0.0)
        matrix = (matrix + sp.diags(self._diag_drop + float(delta), format="csc")).tocsc()
        matrix.sort_indices()
        return matrix

    def _factor_numeric(self, theta, delta):
        """Jacobi scaling + shift ladder + exact factor (PS:839-897)."""
        t0 = time.perf_counter()
        M = self.assemble(self._weights, delta)
        fresh = self._factor is None or not (
            M.indptr.shape == self._pattern[0].shape
            and np.array_equal(M.indptr, self._pattern[0])
            and np.array_equal(M.indices, self._pattern[1]))
        diag = M.diagonal()
        dsc = 1.0 / np.sqrt(np.where(diag > 0.0, diag, 1.0))
        cols = np.repeat(np.arange(self.m), np.diff(M.indptr))
        M.data *= dsc[M.indices] * dsc[cols]
        start = 0
        if self._shift > 0.0:
            start = max(0, int(np.searchsorted(SHIFTS, self._shift, side="left")) - 1)
        last = None
        for eta in SHIFTS[start:]:
            Ms = M if eta == 0.0 else (M + eta * sp.eye(self.m, format="csc")).tocsc()
            try:
                factor = _LibraryFactor(Ms)
            except Exception as exc:
                self.stats["not_pd"] = self.stats.get("not_pd", 0) + 1
                last = exc
                continue
            self._factor = factor
            if fresh:
                self._pattern = (M.indptr.copy(), M.indices.copy())
                self.n_symbolic += 1
            self._shift = eta
            if eta:
                self.n_shifted += 1
            break
        else:
            self.t_fact += time.perf_counter() - t0
            raise KKTSolveError(f"Cholesky failed even with a 1e-2 shift ({last})")
        self._dsc = dsc
        self.theta = theta
        self.delta = delta
        self._prepare_low_rank(theta)
        self.n_fact += 1
        self.t_fact += time.perf_counter() - t0

    def _apply_base(self, r):
        return self._dsc * self._factor.solve(self._dsc * r)

    def _prepare_low_rank(self, theta):
        """Exact Woodbury core for the extracted columns (PS:906-947)."""
        q = self.low_rank_columns.size
        if not q:
            self._lr_basis = self._lr_chol = None
            return
        basis = np.empty((self.m, q), order="F")
        rhs = np.zeros(self.m)
        for j in range(q):
            rhs.fill(0.0)
            s, e = self.A_low_rank.indptr[j], self.A_low_rank.indptr[j + 1]
            rhs[self.A_low_rank.indices[s:e]] = self.A_low_rank.data[s:e]
            basis[:, j] = self._apply_base(rhs)
        core = np.asarray(self.A_low_rank.T @ basis)
        core.flat[:: q + 1] += theta[self.low_rank_columns]
        core = 0.5 * (core + core.T)
        scale = max(float(np.max(np.abs(np.diag(core)))), 1.0)
        last = None
        for rel in (0.0, 1e-14, 1e-12, 1e-10, 1e-8):
            try:
                chol = np.linalg.cholesky(core if rel == 0.0 else core + rel * scale * np.eye(q))
            except np.linalg.LinAlgError as exc:
                last = exc
                continue
            self._lr_basis, self._lr_chol = basis, chol
            return
        raise KKTSolveError(f"low-rank correction is not positive definite ({last})")

    def apply(self, r):
        """z = P^-1 r including the Woodbury correction (PS:949-956)."""
        z = self._apply_base(r)
        if self._lr_chol is None:
            return z
        small = np.asarray(self.A_low_rank.T @ z).ravel()
        coef = np.linalg.solve(self._lr_chol.T, np.linalg.solve(self._lr_chol, small))
        return z - self._lr_basis @ coef

    def full_matvec(self, v):
        """H v = A (W (A' v)) + delta v (SAP:434-438)."""
        return np.asarray(self.A @ (self._weights * (self.At @ v)) + self.delta * v).ravel()

    # ---------------- factorize decisions (SSP:396-467) ----------------
    def _lag_bound(self, weights, delta):
        if self._reference_weights is None:
            return np.inf
        ratio = weights / np.maximum(self._reference_weights, np.finfo(float).tiny)
        fin = ratio[np.isfinite(ratio) & (ratio > 0.0)]
        if fin.size != ratio.size:
            return np.inf
        bound = max(float(fin.max()), 1.0 / float(fin.min()))
        if self._reference_delta > 0.0 and delta > 0.0:
            dr = delta / self._reference_delta
            bound = max(bound, dr, 1.0 / dr)
        elif delta != self._reference_delta:
            return np.inf
        return float(bound)

    def _factor_current(self, theta, delta):
        self._factor_numeric(theta, delta)
        self.factor_age = 0
        self._reference_weights = self._weights.copy()
        self._reference_delta = float(delta)
        self.preconditioner_nnz.append(self._factor.nnz)

    def factorize(self, theta, delta):
        theta = np.asarray(theta, dtype=float)
        self._weights = 1.0 / theta
        spectral_poor = self._last_condition >= self.condition_limit and self._last_cg >= max(10, int(0.8 * self.refresh_cg))
        poor = self._last_cg >= self.refresh_cg or spectral_poor or self._last_stagnated
        self._poor_streak = self._poor_streak + 1 if poor else 0
        can_grow = self.core_target < self.max_core_target
        swap_due = poor and not can_grow and self._poor_streak >= self.swap_patience
        refresh = not self._initialized or self.core_age >= self.core_max_age or (poor and can_grow) or swap_due
        if refresh:
            self._refresh_core(self._weights, delta, stable=bool(self.core.size))
            self._poor_streak = 0
        lag = self._lag_bound(self._weights, float(delta))
        self.lag_spectral_bounds.append(lag)
        if self._factor is None or self.factor_age >= self.factor_max_age or refresh or poor or lag > self.lag_spectral_ratio:
            self._factor_current(theta, float(delta))
        else:
            self.theta = theta
            self.delta = float(delta)
            self.factor_age += 1
        self.core_age += 1
        self.core_sizes.append(int(self.core.size))

    # ---------------- PCG (SSP:469-581) ----------------
    def pcg(self, rhs, tol):
        scale = max(1.0, float(np.linalg.norm(rhs)))
        x = self.apply(rhs)
        r = np.asarray(rhs - self.full_matvec(x)).ravel()
        z = self.apply(r)
        rz = float(r @ z)
        p = z.copy()
        alphas, betas = [], []
        repl = checks = 0
        hist = [float(np.linalg.norm(r)) / scale]
        info = self.cg_maxit
        it = 0
        if hist[-1] <= tol:
            info = 0
        elif not np.isfinite(rz) or rz <= 0.0:
            info = -1
        while info > 0 and it < self.cg_maxit:
            q = self.full_matvec(p)
            curv = float(p @ q)
            floor = np.finfo(float).tiny * max(1.0, float(np.linalg.norm(p)) * float(np.linalg.norm(q)))
            if not np.isfinite(curv) or curv <= floor:
                info = -2
                break
            alpha = rz / curv
            if not np.isfinite(alpha) or alpha <= 0.0:
                info = -3
                break
            x += alpha * p
            r -= alpha * q
            it += 1
            alphas.append(alpha)
            rec = float(np.linalg.norm(r)) / scale
            replace = self.residual_replacement > 0 and it % self.residual_replacement == 0
            if replace or rec <= tol:
                rt = np.asarray(rhs - self.full_matvec(x)).ravel()
                checks += 1
                tn = float(np.linalg.norm(rt))
                rel = tn / scale
                hist.append(rel)
                if not np.isfinite(rel):
                    info = -4
                    break
                if rel <= tol:
                    r = rt
                    info = 0
                    break
                if float(np.linalg.norm(rt - r)) > self.residual_gap_tolerance * max(tn, tol * scale):
                    r = rt
                    repl += 1
                    z = self.apply(r)
                    rz = float(r @ z)
                    if not np.isfinite(rz) or rz <= 0.0:
                        info = -5
                        break
                    p = z.copy()
                    alphas.clear()
                    betas.clear()
                    continue
            z = self.apply(r)
            rz_new = float(r @ z)
            if not np.isfinite(rz_new) or rz_new <= 0.0:
                info = -6
                break
            beta = rz_new / rz
            if not np.isfinite(beta) or beta < 0.0:
                info = -7
                break
            betas.append(beta)
            p = z + beta * p
            rz = rz_new
        r = np.asarray(rhs - self.full_matvec(x)).ravel()
        rel = float(np.linalg.norm(r)) / scale
        if info > 0 and rel <= tol:
            info = 0
        lo, hi, cond = ritz_estimate(alphas, betas)
        wdw = self.stagnation_window
        stagnated = len(hist) >= wdw and hist[-1] > self.stagnation_reduction * hist[-wdw]
        return KrylovResult(x, info, it, rel, float(cond), float(max(abs(lo - 1.0), abs(hi - 1.0))),
                            r, repl, checks, bool(stagnated))

    def _minres(self, r1, r2, tol):
        """Augmented-system MINRES recovery (SSP:583-637)."""
        size = self.n + self.m
        count = [0]

        def matvec(v):
            dx, q = v[: self.n], v[self.n:]
            return np.concatenate([self.theta * dx + self.At @ q, self.A @ dx - self.delta * q])

        def prec(v):
            return np.concatenate([v[: self.n] / self.theta, self.apply(v[self.n:])])

        op = spla.LinearOperator((size, size), matvec=matvec, dtype=float)
        pm = spla.LinearOperator((size, size), matvec=prec, dtype=float)
        rhs = np.concatenate([r1, r2])
        scale = max(1.0, float(np.linalg.norm(rhs)))
        target = max(2.0 * float(tol), self.fail_tol)
        sol = np.zeros_like(rhs)
        for _ in range(3):
            cr = rhs - matvec(sol)
            if np.isfinite(float(np.linalg.norm(cr)) / scale) and float(np.linalg.norm(cr)) / scale <= target:
                break
            corr, info = spla.minres(op, cr, M=pm, rtol=max(np.finfo(float).eps, min(float(tol), 1e-8)),
                                     maxiter=self.minres_maxit, callback=lambda _: count.__setitem__(0, count[0] + 1))
            sol += corr
            if info < 0:
                break
        rel = float(np.linalg.norm(rhs - matvec(sol))) / scale
        return sol[: self.n], -sol[self.n:], bool(np.isfinite(rel) and rel <= target), count[0], rel

    def _record(self, res: KrylovResult):
        self._last_cg = int(res.iterations)
        self._last_condition = float(res.condition_estimate)
        self._last_spectral_error = float(res.spectral_error_estimate)
        self._last_stagnated = bool(res.stagnated)
        self._enrichment = res.residual.copy()
        self.condition_estimates.append(self._last_condition)
        self.spectral_error_estimates.append(self._last_spectral_error)
        self.residual_replacements += int(res.replacements)
        self.residual_checks += int(res.residual_checks)
        self.stagnation_events += int(res.stagnated)

    def solve(self, r1, r2):
        """(dx, dy) of the proximal Newton system (SSP:651-752)."""
        started = time.perf_counter()
        fact_before = self.t_fact
        wr1 = r1 / self.theta
        rhs = np.asarray(r2 - self.A @ wr1).ravel()
        tol = self._linear_tolerance()
        res = self.pcg(rhs, tol)
        total = res.iterations
        failed = res.info != 0 or not np.isfinite(res.relative_residual) or res.relative_residual > tol
        can_grow = self.core_target < self.max_core_target
        if failed and (can_grow or res.info < 0 or res.stagnated) and self.candidates.size and not self.use_full_preconditioner:
            self.pcg_retries += 1
            self._enrichment = res.residual.copy()
            if self.core_target < self.max_core_target:
                self.core_target = min(self.max_core_target, max(self.core_target + 1, int(np.ceil(self.retry_growth * self.core_target))))
            self._refresh_core(self._weights, self.delta, stable=True)
            self._factor_current(self.theta, self.delta)
            self.core_sizes.append(int(self.core.size))
            res = self.pcg(rhs, tol)
            total += res.iterations
            failed = res.info != 0 or not np.isfinite(res.relative_residual) or res.relative_residual > tol
        if failed and res.relative_residual <= max(2.0 * tol, self.fail_tol):
            failed = False
        self.cg_tolerances.append(tol)
        self.cg_iterations.append(total)
        self._record(res)
        if res.info < 0:
            self.krylov_breakdowns += 1
        elif res.info > 0:
            self.krylov_iteration_limits += 1
        if not failed:
            dy = res.solution
            dx = wr1 + (self.At @ dy) / self.theta
        elif self.minres_fallback and self.n + self.m <= self.minres_max_dimension and (res.info < 0 or res.stagnated):
            self.cg_failures += 1
            self.minres_fallbacks += 1
            dx, dy, ok, its, rel = self._minres(r1, r2, tol)
            self.minres_iterations += its
            self.full_fallbacks += 1
            if not ok:
                self.t_solve += max(0.0, time.perf_counter() - started - (self.t_fact - fact_before))
                raise KKTSolveError(f"PCG and MINRES failed: info={res.info}, minres residual={rel:.2e}")
        else:
            self.cg_failures += 1
            self.t_solve += max(0.0, time.perf_counter() - started - (self.t_fact - fact_before))
            raise KKTSolveError(f"PCG failed: info={res.info}, residual={res.relative_residual:.2e}")
        self.n_solve += 1
        self.t_solve += max(0.0, time.perf_counter() - started - (self.t_fact - fact_before))
        return dx, dy
