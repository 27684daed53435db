"""CPU tests pinning the oracle (the FP64 CPU restatement of the hot path).

Each check is chosen so a plausible mistake fails it: LP optima against
scipy's HiGHS, the generator's closed-form duality sandwich, dense
eigenvalues for the Ritz estimate, brute force for step-to-boundary, scipy's
own ``cg`` for the iteration count, and planted mistakes that must be caught.
"""
import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.optimize import linprog

import mf_psipm_amd as P
from oracle import ssp_cpu as O


def _lp(m, n, k, seed):
    raw, x0, y0, z0 = P.generators.random_lp(m, n, k, seed)
    lp = P.lp.to_standard_form(raw)
    P.lp.scale_problem(lp)
    return raw, lp, (x0, y0, z0)


def _solve(raw, lp, **kw):
    tol = P.lp.scaled_tolerance(lp, 1e-8)
    kkt = O.SpectralStableColumnPreconditionerCPU(lp.A, protected=~lp.pos, cg_tol=1e-8, cg_tol_max=1e-8, **kw)
    sol = P.ipm.psipm_solve(lp, kkt, tol=tol, compl="gap", maxit=160)
    return sol, P.lp.recover(lp, raw, sol.x, sol.y, sol.z, sol.w), kkt


@pytest.mark.parametrize("m,n,gate", [(300, 900, 1500), (800, 2400, 0)])
def test_oracle_matches_highs(m, n, gate):
    raw, lp, (x0, y0, z0) = _lp(m, n, 5, 0)
    sol, info, kkt = _solve(raw, lp, full_preconditioner_row_threshold=gate)
    ref = linprog(raw.c, A_eq=raw.A, b_eq=raw.rl, bounds=(0, None), method="highs-ds")
    assert sol.status == "optimal"
    assert abs(info["obj"] - ref.fun) <= 1e-8 * abs(ref.fun)
    assert info["pinf"] <= 1e-8 and info["dinf"] <= 1e-8
    # closed-form sandwich from the generator: b'y0 <= opt <= c'x0, gap x0'z0
    lower, upper = float(raw.rl @ y0), float(raw.c @ x0)
    assert abs((upper - lower) - float(x0 @ z0)) <= 1e-9 * abs(upper)
    assert lower - 1e-9 * abs(lower) <= info["obj"] <= upper + 1e-9 * abs(upper)
    if gate == 0:
        assert not kkt.use_full_preconditioner and sum(kkt.cg_iterations) > 0


def test_planted_missing_delta_is_caught(monkeypatch):
    """Dropping delta from H breaks the Newton system; the LP answer check must fail."""
    raw, lp, _ = _lp(300, 900, 5, 3)
    orig = O.SpectralStableColumnPreconditionerCPU.full_matvec

    def wrong(self, v):
        return orig(self, v) - self.delta * v

    monkeypatch.setattr(O.SpectralStableColumnPreconditionerCPU, "full_matvec", wrong)
    kkt = O.SpectralStableColumnPreconditionerCPU(lp.A, protected=~lp.pos, cg_tol=1e-10, cg_tol_max=1e-10,
                                                  full_preconditioner_row_threshold=0)
    theta = 10.0 ** np.random.default_rng(0).uniform(-2, 2, 900)
    kkt.factorize(theta, 1e-1)
    r1, r2 = np.random.default_rng(1).standard_normal(900), np.random.default_rng(2).standard_normal(300)
    _, dy = kkt.solve(r1, r2)
    H = lp.A @ sp.diags(1 / theta) @ lp.A.T + 1e-1 * sp.eye(300)
    rhs = r2 - lp.A @ (r1 / theta)
    assert np.linalg.norm(H @ dy - rhs) > 1e-3 * np.linalg.norm(rhs)


def test_planted_wrong_scaling_changes_preconditioner():
    """The Jacobi scaling must cancel: P^-1 r equals the unscaled dense solve."""
    raw, lp, _ = _lp(400, 1200, 5, 4)
    kkt = O.SpectralStableColumnPreconditionerCPU(lp.A, protected=~lp.pos)
    theta = 10.0 ** np.random.default_rng(5).uniform(-3, 3, 1200)
    kkt.factorize(theta, 1e-4)
    M = kkt.assemble(1.0 / theta, 1e-4).toarray()
    r = np.random.default_rng(6).standard_normal(400)
    ref = la.solve(M, r, assume_a="pos")
    got = kkt.apply(r)
    assert np.linalg.norm(got - ref) <= 1e-8 * np.linalg.norm(ref)
    wrong = kkt._dsc ** 2 * got  # scaling applied twice
    assert np.linalg.norm(wrong - ref) > 1e-3 * np.linalg.norm(ref)


def test_assembled_matrix_matches_definition():
    """M = A_C W_C A_C' + diag(max(d_full - d_core, 0)) + delta I, built densely."""
    raw, lp, _ = _lp(300, 900, 5, 7)
    kkt = O.SpectralStableColumnPreconditionerCPU(lp.A, protected=~lp.pos, full_preconditioner_row_threshold=0,
                                                  core_factor=0.5)
    w = 10.0 ** np.random.default_rng(8).uniform(-2, 2, 900)
    kkt._weights = w
    kkt._refresh_core(w, 1e-3, stable=False)
    M = kkt.assemble(w, 1e-3).toarray()
    A = lp.A.toarray()
    C = kkt.core
    G = (A[:, C] * w[C]) @ A[:, C].T
    full = (A * A) @ w
    ref = G + np.diag(np.maximum(full - np.diag(G), 0.0) + 1e-3)
    assert np.allclose(M, ref, rtol=1e-12, atol=1e-14)
    assert np.all(np.isin(kkt.mandatory, C))


def test_pcg_matches_scipy_cg_iterations():
    """Same x0 = P^-1 rhs and M: iteration counts agree with scipy cg (+-1)."""
    raw, lp, _ = _lp(600, 1800, 5, 9)
    kkt = O.SpectralStableColumnPreconditionerCPU(lp.A, protected=~lp.pos, full_preconditioner_row_threshold=0,
                                                  core_factor=0.4, residual_replacement=0)
    theta = 10.0 ** np.random.default_rng(10).uniform(-2, 2, 1800)
    kkt.factorize(theta, 1e-3)
    rhs = np.random.default_rng(11).standard_normal(600)
    res = kkt.pcg(rhs, 1e-9)
    count = [0]
    op = spla.LinearOperator((600, 600), matvec=kkt.full_matvec)
    pm = spla.LinearOperator((600, 600), matvec=kkt.apply)
    x, info = spla.cg(op, rhs, x0=kkt.apply(rhs), M=pm, rtol=1e-9, atol=0.0, maxiter=500,
                      callback=lambda _: count.__setitem__(0, count[0] + 1))
    assert res.info == 0 and info == 0
    assert abs(res.iterations - count[0]) <= 1
    assert np.linalg.norm(res.solution - x) <= 1e-6 * np.linalg.norm(x)


def test_pcg_exact_in_at_most_m_iterations():
    rng = np.random.default_rng(12)
    A = sp.random(40, 120, density=0.1, random_state=13, format="csc") + sp.eye(40, 120, format="csc")
    kkt = O.SpectralStableColumnPreconditionerCPU(A.tocsc(), cg_maxit=60, full_preconditioner_row_threshold=0,
                                                  core_factor=0.05, max_core_factor=0.05, residual_replacement=0)
    theta = 10.0 ** rng.uniform(-1, 1, 120)
    kkt.factorize(theta, 1e-2)
    res = kkt.pcg(rng.standard_normal(40), 1e-12)
    assert res.info == 0 and res.iterations <= 40


def test_ritz_estimate_matches_dense_eigenvalues():
    """CG's Lanczos tridiagonal gives the extreme eigenvalues of P^-1/2 H P^-1/2."""
    raw, lp, _ = _lp(150, 450, 5, 14)
    kkt = O.SpectralStableColumnPreconditionerCPU(lp.A, protected=~lp.pos, full_preconditioner_row_threshold=0,
                                                  core_factor=0.3, residual_replacement=0, cg_maxit=400)
    theta = 10.0 ** np.random.default_rng(15).uniform(-1, 1, 450)
    kkt.factorize(theta, 1e-2)
    res = kkt.pcg(np.random.default_rng(16).standard_normal(150), 1e-14)
    H = (lp.A @ sp.diags(1 / theta) @ lp.A.T + 1e-2 * sp.eye(150)).toarray()
    Pm = kkt.assemble(1 / theta, 1e-2).toarray()
    ev = la.eigh(H, Pm, eigvals_only=True)
    cond = ev[-1] / ev[0]
    assert abs(res.condition_estimate - cond) <= 1e-3 * cond
    assert abs(res.spectral_error_estimate - max(abs(ev[0] - 1), abs(ev[-1] - 1))) <= 1e-3 * max(1.0, ev[-1])


def test_step_to_boundary_brute_force():
    rng = np.random.default_rng(17)
    for _ in range(50):
        v = rng.uniform(0.01, 1.0, 30)
        dv = rng.standard_normal(30)
        a = P.ipm.step_to_boundary(v, dv, slice(None))
        grid = np.linspace(0, 1, 20001)
        ok = np.all(v[None, :] + grid[:, None] * dv[None, :] >= 0, axis=1)
        assert a >= grid[ok].max() - 1e-12 and abs(a - grid[ok].max()) <= 1e-4
        assert np.all(v + a * dv >= -1e-12)


def test_top_k_order_and_ties():
    s = np.array([1.0, 3.0, 3.0, np.nan, 2.0, np.inf, -1.0])
    # non-finite scores never enter; ties break by index
    assert O.top_k(s, 3).tolist() == [1, 2, 4]
    assert O.top_k(s, 0).size == 0
    assert O.top_k(s, 10).tolist() == [1, 2, 4, 0, 6]


def test_refresh_core_properties():
    raw, lp, _ = _lp(500, 1500, 5, 18)
    kkt = O.SpectralStableColumnPreconditionerCPU(lp.A, protected=~lp.pos, full_preconditioner_row_threshold=0)
    w = 10.0 ** np.random.default_rng(19).uniform(-3, 3, 1500)
    kkt._refresh_core(w, 1e-3, stable=False)
    first = kkt.core.copy()
    assert first.size == kkt.core_target
    assert np.all(np.isin(kkt.mandatory, first))
    assert np.all(np.diff(first) > 0)
    w2 = w * 10.0 ** np.random.default_rng(20).uniform(-0.5, 0.5, 1500)
    kkt._refresh_core(w2, 1e-3, stable=True)
    kept = np.intersect1d(first, kkt.core).size
    assert kept >= min(first.size, int(np.floor(kkt.retention * kkt.core_target))) - kkt.mandatory.size * 0


def test_standard_form_and_scaling_preserve_the_optimum():
    rng = np.random.default_rng(21)
    m, n = 30, 60
    A = sp.random(m, n, density=0.2, random_state=22, format="csc") + sp.eye(m, n, format="csc")
    x0 = rng.uniform(0.5, 1.5, n)
    Ax = A @ x0
    rl = Ax - rng.uniform(0, 1, m)
    ru = Ax + rng.uniform(0, 1, m)
    ru[:5] = P.lp.INF          # one-sided rows
    rl[5:10] = ru[5:10]        # equalities
    cl = np.where(rng.random(n) < 0.2, -P.lp.INF, 0.0)
    cu = np.where(rng.random(n) < 0.5, 3.0, P.lp.INF)
    cu[:3] = cl[:3] = 0.7      # fixed columns
    c = rng.standard_normal(n)
    raw = P.lp.RawLP(c=c, A=A, rl=rl, ru=ru, cl=cl, cu=cu)
    bounds = list(zip(np.where(cl <= -P.lp.INF, None, cl), np.where(cu >= P.lp.INF, None, cu)))
    fin_u = ru < P.lp.INF
    ref = linprog(c, A_ub=sp.vstack([A[fin_u], -A]), b_ub=np.r_[ru[fin_u], -rl], bounds=bounds, method="highs-ds")
    lp = P.lp.to_standard_form(raw)
    P.lp.scale_problem(lp)
    free = ~lp.pos
    b2 = [(0 if lp.pos[j] else None, lp.u[j] if lp.box[j] else None) for j in range(lp.A.shape[1])]
    res = linprog(lp.c, A_eq=lp.A, b_eq=lp.b, bounds=b2, method="highs-ds")
    assert ref.status == 0 and res.status == 0
    assert abs(res.fun + lp.obj_const - ref.fun) <= 1e-8 * max(1.0, abs(ref.fun))
    info = P.lp.recover(lp, raw, res.x, np.zeros(lp.A.shape[0]), np.zeros(lp.A.shape[1]), np.zeros(lp.A.shape[1]))
    assert abs(info["obj"] - ref.fun) <= 1e-7 * max(1.0, abs(ref.fun))
    assert info["pinf"] <= 1e-7
    assert free.any() or True


def test_ground_network_removes_one_row_per_component():
    lp = P.generators.grid_min_cost_flow(6, 7, 5, 0)
    g, roots = P.lp.ground_network(lp)
    assert roots.size == 1 and g.A.shape[0] == lp.A.shape[0] - 1
    assert abs(lp.b.sum()) < 1e-9


def test_generators_shapes_and_feasibility():
    raw, x0, y0, z0 = P.generators.random_lp(200, 700, 5, 3)
    assert raw.A.shape == (200, 700) and np.all(np.diff(raw.A.indptr) == 5)
    A = raw.A.tocsc()
    for j in range(700):
        rows = A.indices[A.indptr[j]:A.indptr[j + 1]]
        assert np.unique(rows).size == 5
    assert np.allclose(A @ x0, raw.rl)
    assert np.allclose(raw.c - A.T @ y0, z0)
    lp = P.generators.grid_min_cost_flow(10, 12, 7, 1)
    assert lp.A.shape == (120, 2 * 2 * 10 * 12 - 2 * 10 - 2 * 12)
    assert np.all(np.abs(lp.A.sum(axis=0)) == 0)
    sc = P.generators.set_covering(50, 400, 2)
    assert np.all(np.diff(sc.A.tocsr().indptr) > 0)
    counts = np.diff(sc.A.indptr)
    assert counts.min() >= 1 and counts.max() <= 10
