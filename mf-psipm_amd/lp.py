"""LP containers and the one-time pipeline around the hot path.

Host-side preprocessing that turns a general LP into the canonical form the
interior-point driver works on, scales it, and maps a solution back:

* ``RawLP`` / ``StandardLP``   -- psipm_active_pool.py:87-142
* ``to_standard_form``         -- psipm_active_pool.py:340-447
* ``scale_problem``            -- psipm_spectral_preconditioner.py:52-102
  (bounded two-sided infinity-norm equilibration, <= 12 sweeps)
* ``scaled_tolerance``         -- psipm_stable_preconditioner.py:63-94
* ``recover``                  -- psipm_active_pool.py:4310-4344
* ``standard_kkt_metrics``     -- psipm_active_pool.py:3635-3648
* ``ground_network``           -- batch_test_vlsi_spectral.py:242-287

Everything here is FP64 numpy/scipy and runs once per LP.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csgraph

INF = 1.0e20  # "infinite" bound, as in MPS/HiGHS data


@dataclass
class RawLP:
    """min/max c'x + offset  s.t.  rl <= A x <= ru,  cl <= x <= cu."""

    c: np.ndarray
    A: sp.csc_matrix
    rl: np.ndarray
    ru: np.ndarray
    cl: np.ndarray
    cu: np.ndarray
    offset: float = 0.0
    maximize: bool = False
    name: str = ""

    @property
    def shape(self):
        return self.A.shape


@dataclass
class StandardLP:
    """min c'v + obj_const  s.t.  A v = b,  0 <= v_j (j in pos),  v_j <= u_j (j in box)."""

    c: np.ndarray
    A: sp.csc_matrix
    b: np.ndarray
    u: np.ndarray
    pos: np.ndarray
    box: np.ndarray
    cert_u: Optional[np.ndarray] = None
    obj_const: float = 0.0
    shift: Optional[np.ndarray] = None
    sign: Optional[np.ndarray] = None
    scale: Optional[np.ndarray] = None
    row_scale: Optional[np.ndarray] = None
    keep: Optional[np.ndarray] = None
    n_aug: int = 0
    n_orig: int = 0
    rows_keep: Optional[np.ndarray] = None
    m_orig: int = 0


def to_standard_form(raw: RawLP) -> StandardLP:
    """Canonical ``min c'v : A v = b, 0 <= v <= u`` (psipm_active_pool.py:340-447).

    Inequality rows get a ``-1`` slack column bounded by the row range; free
    rows are dropped; fixed columns are substituted out; finite lower bounds
    are shifted to zero, upper-only columns are negated, two-sided columns
    keep their width as a box bound.  (The implied slack caps of
    psipm_active_pool.py:293-337 feed only the out-of-scope pool certificates
    and are not carried.)
    """
    A = raw.A.tocsc()
    c = np.asarray(raw.c, dtype=float)
    m0, n0 = A.shape
    rl, ru = np.asarray(raw.rl, float), np.asarray(raw.ru, float)
    cl, cu = np.asarray(raw.cl, float), np.asarray(raw.cu, float)

    free_row = (rl <= -INF) & (ru >= INF)
    rows_keep = np.flatnonzero(~free_row)
    if free_row.any():
        A = A[rows_keep, :]
        rl, ru = rl[rows_keep], ru[rows_keep]
    m = A.shape[0]

    equality = rl == ru
    slack_rows = np.flatnonzero(~equality)
    n_slack = slack_rows.size
    if n_slack:
        slack = sp.csc_matrix(
            (-np.ones(n_slack), (slack_rows, np.arange(n_slack))), shape=(m, n_slack)
        )
        A = sp.hstack([A, slack], format="csc")
        c = np.concatenate([c, np.zeros(n_slack)])
        lo = np.concatenate([cl, rl[slack_rows]])
        hi = np.concatenate([cu, ru[slack_rows]])
    else:
        lo, hi = cl.copy(), cu.copy()
    b = np.where(equality, rl, 0.0)
    n_aug = A.shape[1]
    lo = np.where(lo <= -INF, -np.inf, lo)
    hi = np.where(hi >= INF, np.inf, hi)

    fixed = lo == hi
    has_lo, has_hi = np.isfinite(lo), np.isfinite(hi)
    upper_only = ~has_lo & has_hi
    shift = np.zeros(n_aug)
    shift[has_lo] = lo[has_lo]
    shift[upper_only] = hi[upper_only]
    shift[fixed] = lo[fixed]
    sign = np.where(upper_only, -1.0, 1.0)

    if np.any(shift != 0.0):
        b = b - A @ shift
    obj_const = float(c @ shift)
    if upper_only.any():
        A = (A @ sp.diags(sign)).tocsc()
        c = c * sign

    u = np.where(has_lo & has_hi, hi - lo, np.inf)
    pos = has_lo | has_hi
    keep = np.flatnonzero(~fixed)
    if fixed.any():
        A = A[:, keep].tocsc()
        c, u, pos = c[keep], u[keep], pos[keep]
    A.sort_indices()
    return StandardLP(
        c=np.ascontiguousarray(c),
        A=A,
        b=np.ascontiguousarray(b),
        u=u,
        pos=pos,
        box=pos & np.isfinite(u),
        cert_u=u.copy(),
        obj_const=obj_const + raw.offset * (-1.0 if raw.maximize else 1.0),
        shift=shift,
        sign=sign,
        scale=np.ones(keep.size),
        row_scale=np.ones(m),
        keep=keep,
        n_aug=n_aug,
        n_orig=n0,
        rows_keep=rows_keep,
        m_orig=m0,
    )


def scale_problem(lp: StandardLP) -> StandardLP:
    """Bounded two-sided infinity-norm equilibration, in place.

    psipm_spectral_preconditioner.py:52-102: up to 12 sweeps; each sweep
    divides rows and columns by the square root of their max-abs entry (step
    clipped to [1e-4, 1e4]); cumulative factors are clipped to [1e-12, 1e12];
    stop once every log row/column max-norm is within 0.05 of zero.
    """
    A = lp.A.tocsc(copy=True)
    m, n = A.shape
    left, right = np.ones(m), np.ones(n)
    tiny = np.finfo(float).tiny
    for _ in range(12):
        absolute = abs(A)
        row_max = np.asarray(absolute.max(axis=1).todense()).ravel()
        col_max = np.asarray(absolute.max(axis=0).todense()).ravel()
        row_max[row_max == 0.0] = 1.0
        col_max[col_max == 0.0] = 1.0
        error = max(
            float(np.max(np.abs(np.log(np.maximum(row_max, tiny))), initial=0.0)),
            float(np.max(np.abs(np.log(np.maximum(col_max, tiny))), initial=0.0)),
        )
        if error <= 0.05:
            break
        new_left = np.clip(left * np.clip(1.0 / np.sqrt(row_max), 1e-4, 1e4), 1e-12, 1e12)
        new_right = np.clip(right * np.clip(1.0 / np.sqrt(col_max), 1e-4, 1e4), 1e-12, 1e12)
        A = sp.diags(new_left / left) @ A @ sp.diags(new_right / right)
        left, right = new_left, new_right
    lp.A = A.tocsc()
    lp.A.sort_indices()
    lp.b = lp.b * left
    lp.c = lp.c * right
    lp.u = lp.u / right
    if lp.cert_u is not None:
        lp.cert_u = lp.cert_u / right
    lp.scale = (np.ones(n) if lp.scale is None else lp.scale) * right
    lp.row_scale = left if lp.row_scale is None else lp.row_scale * left
    return lp


def scaled_tolerance(lp: StandardLP, requested: float, max_tightening: float = 10.0) -> float:
    """Internal tolerance guarding against scaling amplification (SAP:63-94)."""
    requested = float(requested)
    if requested <= 0.0:
        raise ValueError("tolerance must be positive")
    col = np.asarray(lp.scale, float)
    row = np.asarray(lp.row_scale, float)
    dual = (
        (1.0 + float(np.abs(lp.c).max(initial=0.0)))
        * float((1.0 / col).max(initial=1.0))
        / (1.0 + float(np.abs(lp.c / col).max(initial=0.0)))
    )
    primal = (
        (1.0 + float(np.abs(lp.b).max(initial=0.0)))
        * float((1.0 / row).max(initial=1.0))
        / (1.0 + float(np.abs(lp.b / row).max(initial=0.0)))
    )
    return requested / min(max(float(max_tightening), 1.0), max(1.0, dual, primal))


def recover(lp: StandardLP, raw: RawLP, x_std, y_std, z_std, w_std) -> dict:
    """Original-space solution and KKT error (psipm_active_pool.py:4310-4344)."""
    v = np.zeros(lp.n_aug)
    v[lp.keep] = lp.scale * x_std
    x_aug = lp.shift + lp.sign * v
    x = x_aug[: lp.n_orig]
    y = np.zeros(raw.A.shape[0])
    y[lp.rows_keep] = (lp.row_scale if lp.row_scale is not None else 1.0) * y_std
    reduced = np.zeros(lp.n_aug)
    reduced[lp.keep] = (z_std - w_std) / lp.scale
    rc = (reduced / lp.sign)[: lp.n_orig]
    obj = float(raw.c @ x) + raw.offset * (-1.0 if raw.maximize else 1.0)
    if raw.maximize:
        obj = -obj
    Ax = raw.A @ x
    pinf = max(
        float(np.max(np.maximum(raw.rl - Ax, Ax - raw.ru), initial=0.0)),
        float(np.max(np.maximum(raw.cl - x, x - raw.cu), initial=0.0)),
    )
    finite = np.concatenate([raw.rl[np.abs(raw.rl) < INF], raw.ru[np.abs(raw.ru) < INF]])
    pinf /= 1.0 + float(np.abs(finite).max(initial=0.0))
    dres = raw.c - raw.A.T @ y - rc
    fixed = np.ones(lp.n_aug, dtype=bool)
    fixed[lp.keep] = False
    dres[fixed[: lp.n_orig]] = 0.0
    dinf = float(np.abs(dres).max(initial=0.0)) / (1.0 + float(np.abs(raw.c).max(initial=0.0)))
    return dict(x=x, y=y, rc=rc, obj=obj, pinf=pinf, dinf=dinf)


def standard_kkt_metrics(lp: StandardLP, x, y, z, w):
    """Unregularised KKT error in solver scaling (psipm_active_pool.py:3635-3648)."""
    rp = lp.b - lp.A @ x
    rd = lp.c - lp.A.T @ y - z + w
    pinf = float(np.abs(rp).max(initial=0.0)) / (1.0 + float(np.abs(lp.b).max(initial=0.0)))
    dinf = float(np.abs(rd).max(initial=0.0)) / (1.0 + float(np.abs(lp.c).max(initial=0.0)))
    primal = float(lp.c @ x) + lp.obj_const
    dual = float(lp.b @ y) + lp.obj_const
    if lp.box.any():
        dual -= float(lp.u[lp.box] @ w[lp.box])
    return pinf, dinf, abs(primal - dual) / (1.0 + abs(primal)), primal


def ground_network(lp: StandardLP):
    """Drop one redundant balance row per connected component.

    batch_test_vlsi_spectral.py:242-287: in each component the row of
    maximum degree (lowest index on ties) is removed.
    """
    A = lp.A.tocsc()
    counts = np.diff(A.indptr)
    if counts.size and not np.all(counts == 2):
        raise ValueError("network grounding requires two endpoints per arc")
    tail = A.indices[A.indptr[:-1]]
    head = A.indices[A.indptr[:-1] + 1]
    graph = sp.csr_matrix(
        (np.ones(2 * tail.size, dtype=np.int8), (np.r_[tail, head], np.r_[head, tail])),
        shape=(A.shape[0], A.shape[0]),
    )
    ncomp, label = csgraph.connected_components(graph, directed=False, return_labels=True)
    degree = np.bincount(A.indices, minlength=A.shape[0])
    best = np.zeros(ncomp, dtype=degree.dtype)
    np.maximum.at(best, label, degree)
    roots = np.full(ncomp, A.shape[0], dtype=np.int64)
    cand = np.flatnonzero(degree == best[label])
    np.minimum.at(roots, label[cand], cand)
    if np.any(np.abs(lp.b[roots]) > 0.0):
        balance = np.bincount(label, weights=lp.b)
        if np.max(np.abs(balance), initial=0.0) > 1e-12:
            raise ValueError("cannot ground an unbalanced network component")
    keep = np.ones(A.shape[0], dtype=bool)
    keep[roots] = False
    grounded = StandardLP(
        c=lp.c, A=A[keep, :].tocsc(), b=lp.b[keep], u=lp.u, pos=lp.pos, box=lp.box,
        cert_u=lp.cert_u, obj_const=lp.obj_const, scale=lp.scale,
        row_scale=np.ones(int(keep.sum())),
    )
    return grounded, roots
