"""CPU tests of the reduced HSD Newton system: the oracle's reduced
elimination against the dense bordered system it eliminates, and the
product's direction (with an exact stand-in for the GPU KKT object) against
the oracle."""
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import mf_psipm_amd as P
from mf_psipm_amd.hsd import HSDEmbedding, HSDNewtonSystem, HSDState
from oracle import hsd_cpu as O


def _problem(m, n, seed, boxed_frac=0.3):
    rng = np.random.default_rng(seed)
    raw, *_ = P.generators.random_lp(m, n, 4, seed)
    A = raw.A.tocsc()
    c = rng.standard_normal(n)
    b = rng.standard_normal(m)
    upper = np.where(rng.uniform(size=n) < boxed_frac, rng.uniform(1.0, 3.0, n), np.inf)
    D = 10.0 ** rng.uniform(-2, 2, n)
    tau, kappa = 1.3, 0.7
    x = rng.uniform(0.1, 0.9, n) * np.where(np.isfinite(upper), upper * tau, 1.0)
    w = np.where(np.isfinite(upper), rng.uniform(0.1, 1.0, n), 0.0)
    state = SimpleNamespace(x=x, z=rng.uniform(0.1, 1.0, n), w=w, tau=tau, kappa=kappa)
    emb = SimpleNamespace(bar_b=rng.standard_normal(m), bar_c=rng.standard_normal(n), bar_g=0.4)
    rhs = (rng.standard_normal(n), rng.standard_normal(m), 0.3, -0.2, rng.standard_normal(n),
           rng.standard_normal(n), 0.1)
    return A, c, upper, b, D, state, emb, rhs


@pytest.mark.parametrize("seed", [0, 1])
def test_reduced_elimination_equals_bordered_system(seed):
    A, c, upper, b, D, st, emb, rhs = _problem(60, 150, seed)
    args = (A, c, upper, b, D, 1e-6, 1e-4, st, emb, 1e-8, rhs)
    dx, dy, dz, dw, dtau, dtheta, dkappa = O.direction(*args)
    ry, rtau, rtheta = O.bordered_solve(*args)
    assert np.linalg.norm(dy - ry) <= 1e-9 * np.linalg.norm(ry)
    assert abs(dtau - rtau) <= 1e-9 * max(1.0, abs(rtau))
    assert abs(dtheta - rtheta) <= 1e-9 * max(1.0, abs(rtheta))


class ExactKKT:
    """CPU stand-in for the GPU KKT object: exact LU solves of H."""
    error_class = RuntimeError

    def __init__(self, A):
        self.A = A
        self.cg_tol = self.cg_tol_max = self.current_cg_tol = 1e-8
        self.solves = 0

    def factorize(self, theta, delta):
        m = self.A.shape[0]
        self.theta = theta
        self.lu = spla.splu((self.A @ sp.diags(1.0 / theta) @ self.A.T + delta * sp.identity(m)).tocsc())

    def solve(self, r1, r2):
        self.solves += 1
        dy = self.lu.solve(r2 - self.A @ (r1 / self.theta))
        return (r1 + self.A.T @ dy) / self.theta, dy

    def solve_pair(self, r1a, r2a, r1b, r2b):
        return self.solve(r1a, r2a), self.solve(r1b, r2b)


@pytest.mark.parametrize("pair", [True, False])
def test_product_direction_matches_oracle(pair):
    A, c, upper, b, D, st, emb, rhs = _problem(50, 120, 3)
    kkt = ExactKKT(A)
    sys_ = HSDNewtonSystem(kkt, A, c, upper, b, D, 1e-6, 1e-4, HSDState(st.x, st.z, st.w, st.tau, st.kappa),
                           HSDEmbedding(emb.bar_b, emb.bar_c, emb.bar_g), zeta=1e-8, pair=pair)
    got = sys_.direction(*rhs)
    ref = O.direction(A, c, upper, b, D, 1e-6, 1e-4, st, emb, 1e-8, rhs)
    assert got[-1]
    for g, r in zip(got[:4], ref[:4]):
        assert np.linalg.norm(g - r) <= 1e-10 * max(1.0, np.linalg.norm(r))
    for g, r in zip(got[4:7], ref[4:7]):
        assert abs(g - r) <= 1e-10 * max(1.0, abs(r))
    assert sys_.solves == 3 and sys_.pair_solves == (1 if pair else 0)
    # a corrector direction on the same matrix reuses h and the theta row
    rhs2 = tuple(v * 0.5 if isinstance(v, float) else v[::-1].copy() for v in rhs)
    got2 = sys_.direction(*rhs2)
    ref2 = O.direction(A, c, upper, b, D, 1e-6, 1e-4, st, emb, 1e-8, rhs2)
    assert sys_.solves == 4
    assert np.linalg.norm(got2[1] - ref2[1]) <= 1e-10 * np.linalg.norm(ref2[1])
