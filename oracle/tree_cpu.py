"""ORACLE -- test infrastructure only.

CPU FP64 restatement of the reference's network specialisation: the spanning
forest of an incidence matrix and the tree-plus-diagonal preconditioner
factored and solved level by level.  Only ``tests/`` may import it, as the
checker of the GPU tree preconditioner.

Restated (numpy/scipy, no import of the reference) from psipm_v3_hsd.py:

* :397-434   ``NetworkTopology`` fields, depth order and level offsets
* :436-557   ``NetworkTopology.from_incidence`` -- two entries (+1, -1) per
  column, root = row of maximum degree, unweighted BFS tree
  (``csgraph.shortest_path``), one virtual root over several components,
  ``tree_child`` per arc and one protected arc per tree relation
* :653-701   ``TreePreconditioner.factor`` / ``solve``: diagonal of
  A_C W A_C' + delta I, tree-edge weights, LDL' from the deepest level up,
  pivots clipped at max(delta 1e-8, tiny); forward sweep, division, backward
  sweep

Extension (no counterpart in the reference, whose V3 driver never grounds):
``grounded=True`` accepts columns with one entry -- arcs to a grounded row
removed from the system -- as arcs to a virtual node m; the BFS starts there,
so the rows next to the ground are the roots (depth 0).

Parity pinning: the tree factor has no fill, so its solve equals the sparse
direct solve of the explicit tree-plus-diagonal matrix (``tree_matrix``),
which is itself the entrywise restriction of A_C W A_C' + delta I to the
diagonal and the tree edges (checked in the tests against scipy's product);
the topology is checked by its invariants (every tree edge is an arc, depths
are BFS distances, m - components protected arcs).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
from scipy.sparse import csgraph

TINY = float(np.finfo(np.float64).tiny)


@dataclass
class Topology:
    rows: int
    first: np.ndarray
    second: np.ndarray      # -1 for an arc to the ground
    parent: np.ndarray      # -1 at a root
    depth: np.ndarray
    tree_child: np.ndarray  # per arc: the child row of the tree edge it parallels, or -1
    protected: np.ndarray
    components: int
    grounded: bool

    @property
    def max_depth(self) -> int:
        return int(self.depth.max(initial=0))

    def levels(self):
        order = np.argsort(self.depth, kind="stable")
        counts = np.bincount(self.depth, minlength=self.max_depth + 1)
        start = np.concatenate(([0], np.cumsum(counts)))
        return order, start


def from_incidence(A: sp.csc_matrix, grounded: bool = False) -> Topology:
    """Spanning forest of the incidence matrix (V3:436-557)."""
    A = A.tocsc()
    m, n = A.shape
    cnt = np.diff(A.indptr)
    if grounded:
        if cnt.size and (cnt.min() < 1 or cnt.max() > 2):
            raise ValueError("grounded network columns hold one or two entries")
    elif cnt.size and not np.all(cnt == 2):
        raise ValueError("the network backend requires exactly two entries per column")
    first = A.indices[A.indptr[:-1]].astype(np.int64)
    two = cnt == 2
    second = np.full(n, -1, dtype=np.int64)
    second[two] = A.indices[A.indptr[:-1][two] + 1]
    v1 = A.data[A.indptr[:-1]]
    v2 = np.where(two, A.data[np.minimum(A.indptr[:-1] + 1, A.data.size - 1)], -v1)
    if not (np.allclose(np.abs(v1), 1.0) and np.allclose(v1 + v2, 0.0)):
        raise ValueError("network columns must contain one +1 and one -1")
    if grounded:
        g = m  # virtual ground node
        other = np.where(two, second, g)
        graph = sp.csr_matrix((np.ones(2 * n), (np.r_[first, other], np.r_[other, first])), shape=(m + 1, m + 1))
        dist, pred = csgraph.shortest_path(graph, directed=False, indices=g, unweighted=True,
                                           return_predecessors=True)
        if not np.all(np.isfinite(dist)):
            raise ValueError("a grounded network component has no arc to the ground")
        parent = pred[:m].astype(np.int64)
        parent[parent == g] = -1
        depth = (dist[:m] - 1.0).astype(np.int64)
        components = 1
        # tree relation of a ground-adjacent row: its arc to the ground
        rel_child = np.where(~two & (parent[first] == -1), first, -1)
    else:
        graph = sp.csr_matrix((np.ones(2 * n), (np.r_[first, second], np.r_[second, first])), shape=(m, m))
        degree = np.bincount(first, minlength=m) + np.bincount(second, minlength=m)
        components, labels = csgraph.connected_components(graph, directed=False, return_labels=True)
        if components == 1:
            root = int(np.argmax(degree))
            dist, pred = csgraph.shortest_path(graph, directed=False, indices=root, unweighted=True,
                                               return_predecessors=True)
            parent = pred.astype(np.int64)
            parent[root] = -1
            depth = dist.astype(np.int64)
        else:
            best = np.zeros(components, dtype=degree.dtype)
            np.maximum.at(best, labels, degree)
            roots = np.full(components, m, dtype=np.int64)
            cand = np.flatnonzero(degree == best[labels])
            np.minimum.at(roots, labels[cand], cand)
            link = sp.csr_matrix((np.ones(components), (np.zeros(components, dtype=np.int64), roots)), shape=(1, m))
            aug = sp.vstack((sp.hstack((graph, link.T)), sp.hstack((link, sp.csr_matrix((1, 1)))))).tocsr()
            dist, pred = csgraph.shortest_path(aug, directed=False, indices=m, unweighted=True,
                                               return_predecessors=True)
            parent = pred[:m].astype(np.int64)
            parent[parent == m] = -1
            depth = (dist[:m] - 1.0).astype(np.int64)
        rel_child = np.full(n, -1, dtype=np.int64)
    tree_child = np.full(n, -1, dtype=np.int64)
    t1 = two & (parent[first] == second)
    t2 = two & (second >= 0) & (parent[np.maximum(second, 0)] == first)
    tree_child[t1] = first[t1]
    tree_child[t2] = second[t2]
    relation = np.where(tree_child >= 0, tree_child, rel_child)
    pick = np.full(m, n, dtype=np.int64)
    has = np.flatnonzero(relation >= 0)
    np.minimum.at(pick, relation[has], has)
    protected = np.zeros(n, dtype=bool)
    protected[pick[pick < n]] = True
    return Topology(m, first, second, parent, depth, tree_child, protected, int(components), grounded)


def tree_factor(topo: Topology, A: sp.csc_matrix, weights: np.ndarray, delta: float, active=None):
    """Pivots and multipliers of the tree-plus-diagonal LDL' (V3:667-690).

    The diagonal sums a_ij^2 w_j over the active arcs; the tree edge of a
    row carries a_child,j a_parent,j w_j summed over the active arcs parallel
    to it (-sum w_j for +-1 incidence), which enters as edge = -that sum.
    """
    m = topo.rows
    w = np.asarray(weights, dtype=float)
    act = np.ones(w.size, dtype=bool) if active is None else np.asarray(active, dtype=bool)
    wa = np.where(act, w, 0.0)
    A = A.tocsc()
    A2 = A.multiply(A).tocsc()
    diagonal = delta + A2 @ wa
    tree = topo.tree_child >= 0
    j = np.flatnonzero(tree)
    child = topo.tree_child[j]
    par = topo.parent[child]
    coef = np.asarray(A[child, j]).ravel() * np.asarray(A[par, j]).ravel()
    edge = np.zeros(m)
    np.add.at(edge, child, -coef * wa[j])
    pivot = diagonal.copy()
    mult = np.zeros(m)
    order, start = topo.levels()
    for level in range(topo.max_depth, 0, -1):
        nodes = order[start[level]:start[level + 1]]
        piv = np.maximum(pivot[nodes], TINY)
        mult[nodes] = -edge[nodes] / piv
        np.add.at(pivot, topo.parent[nodes], -(edge[nodes] ** 2) / piv)
    np.maximum(pivot, max(delta * 1.0e-8, TINY), out=pivot)
    return pivot, mult, diagonal, edge


def tree_solve(topo: Topology, pivot, mult, rhs):
    """Forward sweep, division, backward sweep (V3:692-701)."""
    v = np.array(rhs, dtype=float, copy=True)
    order, start = topo.levels()
    for level in range(topo.max_depth, 0, -1):
        nodes = order[start[level]:start[level + 1]]
        np.add.at(v, topo.parent[nodes], -mult[nodes] * v[nodes])
    v /= pivot
    for level in range(1, topo.max_depth + 1):
        nodes = order[start[level]:start[level + 1]]
        v[nodes] -= mult[nodes] * v[topo.parent[nodes]]
    return v


def tree_matrix(topo: Topology, diagonal, edge) -> sp.csr_matrix:
    """The explicit tree-plus-diagonal matrix: diagonal, and -edge on every
    (row, parent) pair."""
    m = topo.rows
    kids = np.flatnonzero(topo.parent >= 0)
    rows = np.r_[np.arange(m), kids, topo.parent[kids]]
    cols = np.r_[np.arange(m), topo.parent[kids], kids]
    vals = np.r_[diagonal, -edge[kids], -edge[kids]]
    return sp.csr_matrix((vals, (rows, cols)), shape=(m, m))
