"""CPU tests of the network tree preconditioner: the oracle pinned by the
sparse direct solve of the explicit tree-plus-diagonal matrix, the host
topology against the oracle's and against BFS invariants, and the
pointer-jumping tables against repeated parent steps."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import mf_psipm_amd as P
from mf_psipm_amd.network import jump_tables, network_topology, tree_arcs
from oracle import tree_cpu as T


def _grid(rows, cols, seed, ground=False):
    lp = P.generators.grid_min_cost_flow(rows, cols, 4, seed)
    if ground:
        lp, _ = P.lp.ground_network(lp)
    return lp.A.tocsc()


def _two_grids(seed):
    a = _grid(5, 6, seed)
    b = _grid(4, 4, seed + 1)
    return sp.block_diag((a, b)).tocsc()


CASES = [("grid", lambda: _grid(12, 9, 0), False), ("grounded", lambda: _grid(20, 15, 1, True), True),
         ("two components", lambda: _two_grids(2), False)]


@pytest.mark.parametrize("name,make,grounded", CASES)
def test_tree_oracle_equals_direct_solve_of_tree_matrix(name, make, grounded):
    A = make()
    m, n = A.shape
    rng = np.random.default_rng(3)
    w = 10.0 ** rng.uniform(-3, 3, n)
    active = rng.uniform(size=n) < 0.7
    delta = 1e-4
    topo = T.from_incidence(A, grounded=grounded)
    pivot, mult, diagonal, edge = T.tree_factor(topo, A, w, delta, active)
    M = T.tree_matrix(topo, diagonal, edge)
    # the tree matrix is A_C W A_C' + delta I restricted to the diagonal and the tree edges
    H = (A @ sp.diags(np.where(active, w, 0.0)) @ A.T + delta * sp.identity(m)).tocsr()
    kids = np.flatnonzero(topo.parent >= 0)
    pattern = sp.csr_matrix((np.ones(m + 2 * kids.size), (np.r_[np.arange(m), kids, topo.parent[kids]],
                                                          np.r_[np.arange(m), topo.parent[kids], kids])),
                            shape=(m, m))
    assert abs(H.multiply(pattern > 0) - M).max() <= 1e-12 * abs(H).max()
    b = rng.standard_normal(m)
    ref = spla.spsolve(M.tocsc(), b)
    got = T.tree_solve(topo, pivot, mult, b)
    assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref)


@pytest.mark.parametrize("name,make,grounded", CASES)
def test_host_topology_matches_oracle_and_bfs_invariants(name, make, grounded):
    A = make()
    topo = network_topology(A, grounded)
    ref = T.from_incidence(A, grounded=grounded)
    assert topo.grounded == grounded
    for f in ("parent", "depth", "tree_child", "protected"):
        assert np.array_equal(getattr(topo, f), getattr(ref, f)), f
    kids = np.flatnonzero(topo.parent >= 0)
    assert np.all(topo.depth[kids] == topo.depth[topo.parent[kids]] + 1)
    assert np.all(topo.depth[topo.parent < 0] == 0)
    pair = topo.second >= 0
    assert np.all(np.abs(topo.depth[topo.first[pair]] - topo.depth[topo.second[pair]]) <= 1)
    t = np.flatnonzero(topo.tree_child >= 0)
    ends = np.sort(np.c_[topo.first[t], topo.second[t]], axis=1)
    edge = np.sort(np.c_[topo.tree_child[t], topo.parent[topo.tree_child[t]]], axis=1)
    assert np.array_equal(ends, edge)
    assert topo.protected.sum() == (topo.rows if grounded else topo.rows - topo.components)
    # arc coefficients: a_child,j a_parent,j = -1 on every incidence column
    ptr, arcs, coef = tree_arcs(topo, A)
    assert ptr[-1] == t.size and np.all(coef == -1.0)


def test_jump_tables_follow_parents():
    rng = np.random.default_rng(5)
    m = 400
    parent = np.full(m, -1, dtype=np.int64)
    for i in range(1, m):  # random forest: parent index below the child
        if rng.uniform() < 0.97:
            parent[i] = rng.integers(0, i)
    depth = np.zeros(m, dtype=np.int64)
    for i in range(m):
        depth[i] = 0 if parent[i] < 0 else depth[parent[i]] + 1
    rounds, jump, ptr, idx = jump_tables(parent, int(depth.max()))
    assert (1 << rounds) > depth.max() and (rounds == 0 or (1 << (rounds - 1)) <= depth.max())

    def up(y, steps):
        for _ in range(steps):
            if y < 0:
                return -1
            y = parent[y]
        return y

    for k in range(rounds):
        for y in range(m):
            assert jump[k, y] == up(y, 1 << k)
        for x in range(m):
            kids = idx[ptr[k, x]:ptr[k, x + 1]]
            assert np.array_equal(kids, np.flatnonzero(jump[k] == x))


@pytest.mark.parametrize("name,make,grounded", CASES)
def test_pointer_jumping_rounds_equal_level_sweeps(name, make, grounded):
    """The device solve's round formulas, evaluated in numpy on the host
    tables, reproduce the oracle's level sweeps."""
    A = make()
    m, n = A.shape
    rng = np.random.default_rng(7)
    w = 10.0 ** rng.uniform(-2, 2, n)
    topo = network_topology(A, grounded)
    ref_topo = T.from_incidence(A, grounded=grounded)
    pivot, mult, _, _ = T.tree_factor(ref_topo, A, w, 1e-3)
    rounds, jump, ptr, idx = jump_tables(topo.parent, topo.max_depth)
    W = np.zeros((max(rounds, 1), m))
    W[0] = np.where(topo.parent >= 0, -mult, 0.0)
    for k in range(1, rounds):
        a = jump[k - 1]
        W[k] = np.where(a >= 0, W[k - 1] * W[k - 1][np.maximum(a, 0)], 0.0)
    b = rng.standard_normal(m)
    s = b.copy()
    for k in range(rounds):
        t = s.copy()
        for x in range(m):
            ys = idx[ptr[k, x]:ptr[k, x + 1]]
            t[x] += np.dot(W[k][ys], s[ys])
        s = t
    s /= pivot
    for k in range(rounds):
        a = jump[k]
        s = s + np.where(a >= 0, W[k] * s[np.maximum(a, 0)], 0.0)
    ref = T.tree_solve(ref_topo, pivot, mult, b)
    assert np.linalg.norm(s - ref) <= 1e-12 * np.linalg.norm(ref)


def test_non_network_matrix_is_rejected():
    A = sp.random(20, 50, density=0.2, random_state=0, format="csc")
    with pytest.raises(ValueError):
        network_topology(A, grounded=False)
    with pytest.raises(ValueError):
        T.from_incidence(A)
