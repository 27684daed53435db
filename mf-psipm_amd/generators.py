"""Synthetic LP generators for the measured workloads.

The reference publishes no generator; these follow the workload definitions
used for measurement (all seeded with ``numpy.random.default_rng``):

* ``random_lp``       -- standard-form ``min c'x : Ax = b, x >= 0`` with k
  nonzeros per column at distinct uniform rows, values N(0,1);
  ``b = A x0`` with ``x0 ~ U[0.1, 1]`` and ``c = A'y0 + z0`` with
  ``y0 ~ N(0,1)``, ``z0 ~ U[0.1, 1]`` (so the LP is feasible and bounded and
  ``b'y0 <= opt <= c'x0`` with ``c'x0 - b'y0 = x0'z0``).
  Sizes: 2,000 x 6,000 x 5/col (CPU reference case), 200,000 x 1,000,000 x
  8/col (single-GPU case), 64 LPs of m ~ U{5,000..50,000}, n = 3m, 6/col.
* ``grid_min_cost_flow`` -- uncapacitated min-cost flow on a directed
  R x C 4-neighbour grid (both directions), integer costs U[1,100], P
  source/sink pairs with integer demands U[1,1000]; node-arc incidence with
  +1 at the tail and -1 at the head (batch_test_vlsi_spectral.py:206-213).
* ``set_covering``    -- ``min c'x : Ax >= 1, x >= 0``, column j covers
  k ~ U{2..10} distinct rows with coefficient 1, costs U{1,2,3}.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .lp import INF, RawLP, StandardLP


def _distinct_rows(rng, m: int, n: int, k: np.ndarray | int) -> tuple:
    """Row indices of n columns, column j holding k_j distinct uniform rows.

    Returns (indptr, indices) of a CSC pattern with sorted rows per column.
    Columns with a repeated row are redrawn until all rows are distinct, which
    gives the uniform distribution over k-subsets.
    """
    counts = np.full(n, int(k), dtype=np.int64) if np.isscalar(k) else np.asarray(k, np.int64)
    kmax = int(counts.max(initial=0))
    rows = rng.integers(0, m, size=(n, kmax))
    valid = np.arange(kmax)[None, :] < counts[:, None]
    sentinel = np.iinfo(np.int64).max
    while True:
        keyed = np.where(valid, rows, sentinel)
        keyed.sort(axis=1)
        dup = np.any((keyed[:, 1:] == keyed[:, :-1]) & (keyed[:, 1:] != sentinel), axis=1)
        if not dup.any():
            rows = keyed
            break
        redraw = np.flatnonzero(dup)
        rows = keyed
        rows[redraw] = rng.integers(0, m, size=(redraw.size, kmax))
    indices = rows[valid]  # sorted rows, sentinels last
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    return indptr, indices.astype(np.int32)


def random_lp(m: int, n: int, nnz_per_col: int, seed: int):
    """Feasible, bounded random standard-form LP; returns (RawLP, x0, y0, z0)."""
    rng = np.random.default_rng(seed)
    indptr, indices = _distinct_rows(rng, m, n, nnz_per_col)
    values = rng.standard_normal(indices.size)
    A = sp.csc_matrix((values, indices, indptr), shape=(m, n))
    x0 = rng.uniform(0.1, 1.0, n)
    y0 = rng.standard_normal(m)
    z0 = rng.uniform(0.1, 1.0, n)
    b = A @ x0
    c = A.T @ y0 + z0
    raw = RawLP(
        c=c, A=A, rl=b.copy(), ru=b.copy(), cl=np.zeros(n), cu=np.full(n, INF),
        name=f"random_{m}x{n}x{nnz_per_col}_s{seed}",
    )
    return raw, x0, y0, z0


def batch_sizes(count: int = 64, seed0: int = 100, m_lo: int = 5000, m_hi: int = 50000):
    """Row counts of the batched workload: LP i uses seed ``seed0 + i``."""
    return [int(np.random.default_rng(seed0 + i).integers(m_lo, m_hi + 1)) for i in range(count)]


def grid_min_cost_flow(rows: int, cols: int, pairs: int, seed: int) -> StandardLP:
    """Uncapacitated min-cost flow on a directed 4-neighbour grid (StandardLP)."""
    rng = np.random.default_rng(seed)
    node = np.arange(rows * cols).reshape(rows, cols)
    right = np.stack([node[:, :-1].ravel(), node[:, 1:].ravel()], axis=1)
    down = np.stack([node[:-1, :].ravel(), node[1:, :].ravel()], axis=1)
    und = np.concatenate([right, down])
    arcs = np.concatenate([und, und[:, ::-1]])  # both directions
    tail, head = arcs[:, 0], arcs[:, 1]
    narc = arcs.shape[0]
    nodes = rows * cols
    costs = rng.integers(1, 101, size=narc).astype(float)
    b = np.zeros(nodes)
    for _ in range(pairs):
        s, t = rng.choice(nodes, size=2, replace=False)
        d = float(rng.integers(1, 1001))
        b[s] += d
        b[t] -= d
    lo = np.minimum(tail, head)
    hi = np.maximum(tail, head)
    indices = np.stack([lo, hi], axis=1).ravel().astype(np.int32)
    values = np.where(tail < head, 1.0, -1.0)
    values = np.stack([values, -values], axis=1).ravel()
    indptr = np.arange(0, 2 * narc + 1, 2, dtype=np.int64)
    A = sp.csc_matrix((values, indices, indptr), shape=(nodes, narc))
    return StandardLP(
        c=costs, A=A, b=b, u=np.full(narc, np.inf), pos=np.ones(narc, bool),
        box=np.zeros(narc, bool), cert_u=None, obj_const=0.0, scale=np.ones(narc),
        row_scale=np.ones(nodes),
    )


def set_covering(m: int, n: int, seed: int, k_lo: int = 2, k_hi: int = 10) -> RawLP:
    """``min c'x : Ax >= 1, x >= 0`` with 0/1 columns covering k~U{k_lo..k_hi} rows."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(k_lo, k_hi + 1, size=n)
    indptr, indices = _distinct_rows(rng, m, n, counts)
    A = sp.csc_matrix((np.ones(indices.size), indices, indptr), shape=(m, n))
    # make every row coverable
    empty = np.flatnonzero(np.diff(A.tocsr().indptr) == 0)
    if empty.size:
        extra = sp.csc_matrix((np.ones(empty.size), (empty, np.arange(empty.size))), shape=(m, empty.size))
        A = sp.hstack([A, extra], format="csc")
        n = A.shape[1]
    c = rng.integers(1, 4, size=n).astype(float)
    return RawLP(
        c=c, A=A.tocsc(), rl=np.ones(m), ru=np.full(m, INF), cl=np.zeros(n),
        cu=np.full(n, INF), name=f"setcover_{m}x{n}_s{seed}",
    )


def mixed_bounds_lp(m: int, n: int, nnz_per_col: int, seed: int) -> RawLP:
    """Random LP exercising every bound and row type of the pipeline.

    Columns: 40% ``[0, inf)``, 30% boxed ``[l, u]`` around an interior point,
    10% free, 10% upper-bounded only, 10% with a negative finite lower bound;
    rows: 70% equalities, 30% ranges around ``A x0``.  The costs make ``y0``
    dual feasible (free columns get exactly ``(A'y0)_j``), so the LP is
    feasible and bounded.
    """
    rng = np.random.default_rng(seed)
    indptr, indices = _distinct_rows(rng, m, n, nnz_per_col)
    A = sp.csc_matrix((rng.standard_normal(indices.size), indices, indptr), shape=(m, n))
    kind = rng.choice(5, size=n, p=[0.4, 0.3, 0.1, 0.1, 0.1])
    x0 = rng.uniform(0.1, 1.0, n)
    x0[kind == 3] = -x0[kind == 3]
    cl = np.zeros(n)
    cu = np.full(n, INF)
    box = kind == 1
    cl[box] = x0[box] - rng.uniform(0.1, 1.0, box.sum())
    cu[box] = x0[box] + rng.uniform(0.1, 1.0, box.sum())
    cl[kind == 2] = -INF
    cl[kind == 3] = -INF
    cu[kind == 3] = 0.0
    neg = kind == 4
    cl[neg] = -rng.uniform(0.5, 2.0, neg.sum())
    y0 = rng.standard_normal(m)
    z0 = rng.uniform(0.1, 1.0, n)
    aty = A.T @ y0
    c = aty + z0
    c[kind == 1] = aty[kind == 1] + rng.standard_normal(box.sum())
    c[kind == 2] = aty[kind == 2]
    c[kind == 3] = aty[kind == 3] - z0[kind == 3]
    ax = A @ x0
    rng_rows = rng.uniform(size=m) < 0.3
    rl, ru = ax.copy(), ax.copy()
    rl[rng_rows] -= rng.uniform(0.1, 1.0, rng_rows.sum())
    ru[rng_rows] += rng.uniform(0.1, 1.0, rng_rows.sum())
    return RawLP(c=c, A=A, rl=rl, ru=ru, cl=cl, cu=cu, name=f"mixed_{m}x{n}_s{seed}")
