"""GPU tests of the network tree-plus-diagonal preconditioner (psg_tree_*):
the device factor and pointer-jumping solve against the oracle's level
sweeps, the PCG on a network Newton system against a sparse direct solve,
and the interior-point method on a grid min-cost flow against HiGHS."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import mf_psipm_amd as P
from mf_psipm_amd.network import HipTreePreconditioner
from oracle import tree_cpu as T

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


def _grid_lp(rows, cols, pairs, seed, ground=True):
    lp = P.generators.grid_min_cost_flow(rows, cols, pairs, seed)
    if ground:
        lp, _ = P.lp.ground_network(lp)
    return lp


def _two_components():
    a = _grid_lp(9, 11, 3, 4, ground=False).A
    b = _grid_lp(7, 6, 3, 5, ground=False).A
    return sp.block_diag((a, b)).tocsc()


@pytest.mark.parametrize("case", ["grounded 30x30", "grid 25x20", "two components", "grounded 3x200"])
def test_tree_factor_and_apply_match_oracle(gpu, case):
    _say(f"tree factor/apply {case}")
    if case == "grounded 30x30":
        A, grounded = _grid_lp(30, 30, 10, 1).A.tocsc(), True
    elif case == "grid 25x20":
        A, grounded = _grid_lp(25, 20, 10, 2, ground=False).A.tocsc(), False
    elif case == "two components":
        A, grounded = _two_components(), False
    else:  # long thin grid: deep tree, several jump rounds
        A, grounded = _grid_lp(3, 200, 4, 3).A.tocsc(), True
    m, n = A.shape
    rng = np.random.default_rng(11)
    theta = 10.0 ** rng.uniform(-4, 4, n)
    delta = 1e-6
    obj = HipTreePreconditioner(A, grounded=grounded)
    obj.factorize(theta, delta)
    piv, mult = obj.tree_factor_values()
    topo = T.from_incidence(A, grounded=grounded)
    rp, rm, _, _ = T.tree_factor(topo, A, 1.0 / theta, delta)
    assert np.abs(piv - rp).max() <= 1e-13 * np.abs(rp).max()
    assert np.abs(mult - rm).max() <= 1e-12 * max(1.0, np.abs(rm).max())
    for s in range(3):
        r = np.random.default_rng(20 + s).standard_normal(m)
        z = obj._apply(r)
        ref = T.tree_solve(topo, rp, rm, r)
        assert np.linalg.norm(z - ref) <= 1e-12 * np.linalg.norm(ref)
        assert np.array_equal(obj._apply(r), z)  # fixed summation order, chunked lists included
    # two interleaved columns through the block apply
    R = np.ascontiguousarray(np.random.default_rng(30).standard_normal((m, 2)))
    Z = np.empty((m, 2))
    obj._call("psg_apply_P2", obj._ctx, R, Z)
    for q in range(2):
        ref = T.tree_solve(topo, rp, rm, R[:, q])
        assert np.linalg.norm(Z[:, q] - ref) <= 1e-12 * np.linalg.norm(ref)
    obj.close()


def test_tree_pcg_solves_network_newton_system(gpu):
    _say("tree PCG network Newton system 60x60")
    lp = _grid_lp(60, 60, 40, 6)
    A = lp.A.tocsc()
    m, n = A.shape
    rng = np.random.default_rng(12)
    theta = 10.0 ** rng.uniform(-2, 2, n)
    delta = 1e-8
    obj = HipTreePreconditioner(A, cg_tol=1e-10, cg_tol_max=1e-10, cg_maxit=3000)
    obj.factorize(theta, delta)
    r1, r2 = rng.standard_normal(n), rng.standard_normal(m)
    dx, dy = obj.solve(r1, r2)
    H = (A @ sp.diags(1.0 / theta) @ A.T + delta * sp.identity(m)).tocsc()
    rhs = r2 - A @ (r1 / theta)
    ref = spla.spsolve(H, rhs)
    assert np.linalg.norm(dy - ref) <= 1e-7 * np.linalg.norm(ref)
    assert np.allclose(dx, (r1 + A.T @ dy) / theta, rtol=1e-9, atol=1e-9 * np.abs(dx).max())
    assert obj.cg_iterations[-1] > 0
    obj.close()


def test_tree_preconditioned_ipm_matches_highs(gpu):
    """Min-cost flow on a 30 x 30 grid through the network pipeline with the
    tree preconditioner (power-of-two cost scaling as
    batch_test_vlsi_spectral.py:340-346)."""
    _say("tree-preconditioned IPM grid 30x30")
    from scipy.optimize import linprog
    full = P.generators.grid_min_cost_flow(30, 30, 40, 7)
    lp, _ = P.lp.ground_network(full)
    c0 = lp.c.copy()
    osc = float(2.0 ** np.ceil(np.log2(np.abs(c0).max())))
    lp.c = c0 / osc
    obj = HipTreePreconditioner(lp.A, protected=~lp.pos, cg_tol=1e-8, cg_tol_max=1e-8, cg_maxit=5000)
    sol = P.ipm.psipm_solve(lp, obj, tol=1e-8, compl="gap", maxit=200)
    ref = linprog(c0, A_eq=lp.A, b_eq=lp.b, bounds=(0, None), method="highs-ds")
    assert sol.status == "optimal"
    assert abs(sol.obj * osc - ref.fun) <= 1e-8 * abs(ref.fun)
    assert np.abs(full.b - full.A @ sol.x).max() <= 1e-8 * (1.0 + np.abs(full.b).max())
    obj.close()
