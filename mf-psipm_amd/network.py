"""Network LPs: spanning forest of an incidence matrix and the GPU
tree-plus-diagonal preconditioner.

Host side of psipm_v3_hsd.py:397-712 (``NetworkTopology``,
``TreePreconditioner``):

* ``network_topology`` builds the spanning forest the way
  ``NetworkTopology.from_incidence`` does (V3:436-557): columns with one +1
  and one -1, root = row of maximum degree (lowest index on ties), the
  unweighted BFS tree of ``csgraph.shortest_path``, one virtual root joining
  several components, ``tree_child`` per arc, one protected arc per tree
  relation.  With ``grounded`` (no counterpart in the reference, whose V3
  driver never grounds; the SSP network pipeline of
  batch_test_vlsi_spectral.py:242-287 does) columns with a single entry are
  arcs to the removed ground rows: they join a virtual node m from which the
  BFS starts, so the rows next to the ground are the roots.
* ``jump_tables`` adds what the device solve needs: the ancestors 2^k levels
  up and, per round, the rows each ancestor collects from.
* ``HipTreePreconditioner`` is the SSP-shaped GPU object (same ``factorize``
  / ``solve`` / counters as ``HipSpectralStableColumnPreconditioner``) whose
  PCG applies the tree factor (``psg_tree_factor``) instead of the
  active-column Cholesky factor.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
from scipy.sparse import csgraph

from .preconditioner import HipSpectralStableColumnPreconditioner


@dataclass
class NetworkTopology:
    rows: int
    first: np.ndarray       # int64: first row of each arc
    second: np.ndarray      # int64: second row, -1 for an arc to the ground
    parent: np.ndarray      # int64: -1 at a root
    depth: np.ndarray       # int64
    tree_child: np.ndarray  # int64 per arc: child row of the tree edge it parallels, or -1
    protected: np.ndarray   # bool per arc: one arc per tree relation
    components: int
    grounded: bool

    @property
    def max_depth(self) -> int:
        return int(self.depth.max(initial=0))

    def depth_order(self):
        """Rows sorted by depth (ascending index within a depth) and the
        offsets of each depth."""
        order = np.argsort(self.depth, kind="stable")
        start = np.zeros(self.max_depth + 2, dtype=np.int64)
        np.cumsum(np.bincount(self.depth, minlength=self.max_depth + 1), out=start[1:])
        return order, start


def _arc_rows(A: sp.csc_matrix, grounded: bool):
    counts = np.diff(A.indptr)
    lo, hi = (1, 2) if grounded else (2, 2)
    if counts.size and (counts.min() < lo or counts.max() > hi):
        raise ValueError("network columns must hold two entries (one with a grounded row removed)")
    start = A.indptr[:-1]
    first = A.indices[start].astype(np.int64)
    pair = counts == 2
    second = np.full(counts.size, -1, dtype=np.int64)
    second[pair] = A.indices[start[pair] + 1]
    a1 = A.data[start]
    a2 = np.where(pair, A.data[np.minimum(start + 1, max(A.data.size - 1, 0))], -a1)
    if not (np.allclose(np.abs(a1), 1.0) and np.allclose(a1 + a2, 0.0)):
        raise ValueError("network columns must contain one +1 and one -1")
    return first, second, pair


def network_topology(A: sp.csc_matrix, grounded=None) -> NetworkTopology:
    """Spanning forest of an incidence matrix (V3:436-557); ``grounded=None``
    detects single-entry columns."""
    A = A.tocsc()
    A.sort_indices()
    m, n = A.shape
    if grounded is None:
        grounded = bool(n and np.diff(A.indptr).min() == 1)
    first, second, pair = _arc_rows(A, grounded)
    rel_child = np.full(n, -1, dtype=np.int64)
    if grounded:
        other = np.where(pair, second, m)
        g = sp.coo_matrix((np.ones(2 * n), (np.r_[first, other], np.r_[other, first])), shape=(m + 1, m + 1))
        dist, pred = csgraph.shortest_path(g.tocsr(), directed=False, indices=m, unweighted=True,
                                           return_predecessors=True)
        if not np.all(np.isfinite(dist)):
            raise ValueError("a grounded network component has no arc to the ground")
        parent = pred[:m].astype(np.int64)
        parent[parent == m] = -1
        depth = dist[:m].astype(np.int64) - 1
        components = 1
        to_ground = ~pair & (parent[first] == -1)
        rel_child[to_ground] = first[to_ground]
    else:
        g = sp.coo_matrix((np.ones(2 * n), (np.r_[first, second], np.r_[second, first])), shape=(m, m)).tocsr()
        degree = np.bincount(first, minlength=m) + np.bincount(second, minlength=m)
        components, label = csgraph.connected_components(g, directed=False, return_labels=True)
        if components == 1:
            root = int(np.argmax(degree))
            dist, pred = csgraph.shortest_path(g, directed=False, indices=root, unweighted=True,
                                               return_predecessors=True)
            parent = pred.astype(np.int64)
            parent[root] = -1
            depth = dist.astype(np.int64)
        else:
            # one virtual root over the component representatives (maximum
            # degree, lowest index), discarded after the BFS
            top = np.zeros(components, dtype=degree.dtype)
            np.maximum.at(top, label, degree)
            reps = np.full(components, m, dtype=np.int64)
            cand = np.flatnonzero(degree == top[label])
            np.minimum.at(reps, label[cand], cand)
            link = sp.coo_matrix((np.ones(components), (reps, np.full(components, m))), shape=(m + 1, m + 1))
            aug = (sp.block_diag((g, sp.csr_matrix((1, 1)))) + link + link.T).tocsr()
            dist, pred = csgraph.shortest_path(aug, directed=False, indices=m, unweighted=True,
                                               return_predecessors=True)
            parent = pred[:m].astype(np.int64)
            parent[parent == m] = -1
            depth = dist[:m].astype(np.int64) - 1
    tree_child = np.full(n, -1, dtype=np.int64)
    up1 = pair & (parent[first] == second)
    up2 = pair & (parent[np.maximum(second, 0)] == first)
    tree_child[up1] = first[up1]
    tree_child[up2] = second[up2]
    relation = np.where(tree_child >= 0, tree_child, rel_child)
    pick = np.full(m, n, dtype=np.int64)
    live = np.flatnonzero(relation >= 0)
    np.minimum.at(pick, relation[live], live)
    protected = np.zeros(n, dtype=bool)
    protected[pick[pick < n]] = True
    expected = m if grounded else m - components
    if int(protected.sum()) != expected:
        raise RuntimeError("failed to construct a spanning-forest column core")
    return NetworkTopology(m, first, second, parent, depth, tree_child, protected, int(components), bool(grounded))


def jump_tables(parent: np.ndarray, max_depth: int):
    """Pointer-jumping tables of the forest: rounds = bit length of
    max_depth; jump[k] = ancestor 2^k levels up (-1 past a root); child_ptr[k]
    / child_idx = for each row x the rows y with jump[k][y] = x, ascending
    (absolute offsets into child_idx)."""
    m = parent.size
    rounds = int(max_depth).bit_length()
    jump = np.empty((rounds, m), dtype=np.int32)
    ptr = np.empty((rounds, m + 1), dtype=np.int64)
    idx = []
    base = 0
    for k in range(rounds):
        if k == 0:
            jump[0] = parent
        else:
            prev = jump[k - 1]
            jump[k] = np.where(prev >= 0, prev[np.maximum(prev, 0)], -1)
        rows = np.flatnonzero(jump[k] >= 0)
        tgt = jump[k][rows]
        kids = rows[np.argsort(tgt, kind="stable")]
        ptr[k, 0] = base
        np.cumsum(np.bincount(tgt, minlength=m), out=ptr[k, 1:])
        ptr[k, 1:] += base
        idx.append(kids.astype(np.int32))
        base += kids.size
    child_idx = np.concatenate(idx) if idx else np.empty(0, dtype=np.int32)
    return rounds, jump, ptr, child_idx


def tree_arcs(topo: NetworkTopology, A: sp.csc_matrix):
    """Per row x, the columns parallel to x's tree edge and a_x,j a_parent,j."""
    A = A.tocsc()
    j = np.flatnonzero(topo.tree_child >= 0)
    child = topo.tree_child[j]
    order = np.lexsort((j, child))
    j, child = j[order], child[order]
    coef = np.asarray(A[child, j]).ravel() * np.asarray(A[topo.parent[child], j]).ravel()
    ptr = np.zeros(topo.rows + 1, dtype=np.int64)
    np.cumsum(np.bincount(child, minlength=topo.rows), out=ptr[1:])
    return ptr, j.astype(np.int32), coef.astype(np.float64)


class HipTreePreconditioner(HipSpectralStableColumnPreconditioner):
    """The proximal Newton solve of a network LP with the GPU PCG and the
    tree-plus-diagonal preconditioner of H = A W A' + delta I (all columns
    active, so the diagonal is H's own; V3:653-701).

    ``factorize(theta, delta)`` refactors the tree every call (one device
    launch sequence, no symbolic work); ``solve`` is the spectral object's
    PCG with its failure ladder minus the core growth, which has no meaning
    here.
    """

    def __init__(self, A: sp.csc_matrix, protected=None, grounded=None, **kw):
        kw.setdefault("structural_backbone", False)
        super().__init__(A, protected=protected, **kw)
        self.topology = network_topology(self.A, grounded)
        topo = self.topology
        order, start = topo.depth_order()
        rounds, jump, cptr, cidx = jump_tables(topo.parent, topo.max_depth)
        aptr, aidx, acoef = tree_arcs(topo, self.A)
        self.tree_rounds = rounds
        self._call("psg_tree_setup", self._ctx, topo.max_depth, np.ascontiguousarray(topo.parent, dtype=np.int32),
                   np.ascontiguousarray(order, dtype=np.int32), start, rounds,
                   np.ascontiguousarray(jump.ravel()) if rounds else None,
                   np.ascontiguousarray(cptr.ravel()) if rounds else None,
                   cidx if cidx.size else None, aptr, aidx if aidx.size else None, acoef if acoef.size else None)
        # no active-column core to grow on a failed solve
        self.candidates = np.empty(0, dtype=np.int64)
        self.core_target = self.max_core_target = 0
        self.tree_edges = int((topo.parent >= 0).sum())

    def factorize(self, theta, delta):
        started = time.perf_counter()
        theta = np.ascontiguousarray(theta, dtype=float)
        self._weights = 1.0 / theta
        self._call("psg_set_weights", self._ctx, theta, float(delta))
        self._call("psg_tree_factor", self._ctx, None)
        self.theta = theta
        self.delta = float(delta)
        self._have_factor = True
        self.n_fact += 1
        self.nnz_factor = self.m + self.tree_edges
        self.preconditioner_nnz.append(self.nnz_factor)
        self.core_sizes.append(0)
        self.t_fact += time.perf_counter() - started

    def tree_factor_values(self):
        """(pivots, multipliers) of the current factor."""
        piv, mult = np.empty(self.m), np.empty(self.m)
        self._call("psg_tree_export", self._ctx, piv, mult)
        return piv, mult
