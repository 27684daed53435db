"""Numpy executor of a supernodal plan (test helper, CPU only).

Runs the same right-looking supernodal algorithm the GPU kernels run --
scatter-map assembly into column-major panels, per-supernode dense POTRF and
TRSM, then the update L_b L_b' scattered into ancestor panels through the
row structures -- so the plan's index data can be checked on the CPU against
a dense Cholesky of P M P'.
"""
import numpy as np


def assemble_panels(plan, A, w, diag, eta=0.0):
    """Panels of D P M P' D + eta I with D = diag(M)^-1/2 (lower part only)."""
    panels = np.zeros(int(plan["poff"][-1]))
    vals = A.data
    seg_ptr, tgt = plan["seg_ptr"], plan["seg_target"]
    contrib = w[plan["pair_col"]] * vals[plan["pair_a"]] * vals[plan["pair_b"]]
    sums = np.add.reduceat(contrib, seg_ptr[:-1]) if contrib.size else np.zeros(0)
    perm = plan["perm"]
    dsc = 1.0 / np.sqrt(diag)
    # target slot -> (row, col) in permuted numbering
    rows_of_slot, cols_of_slot = slot_coordinates(plan)
    r = rows_of_slot[tgt]
    c = cols_of_slot[tgt]
    panels[tgt] = sums * dsc[perm[r]] * dsc[perm[c]]
    panels[plan["diag_index"]] = diag[perm] * dsc[perm] * dsc[perm] + eta
    return panels


def slot_coordinates(plan):
    n = int(plan["poff"][-1])
    rr = np.empty(n, np.int64)
    cc = np.empty(n, np.int64)
    sf, rptr, rows, poff = plan["sfirst"], plan["rptr"], plan["rows"], plan["poff"]
    for s in range(len(sf) - 1):
        w = sf[s + 1] - sf[s]
        R = rows[rptr[s]:rptr[s + 1]]
        nr = R.size
        blk = slice(poff[s], poff[s + 1])
        rr[blk] = np.tile(R, w)
        cc[blk] = np.repeat(np.arange(sf[s], sf[s] + w), nr)
    return rr, cc


def factor_panels(plan, panels):
    """Right-looking supernodal Cholesky in place; returns first failing column or -1."""
    sf, rptr, rows, poff = plan["sfirst"], plan["rptr"], plan["rows"], plan["poff"]
    snode_of = np.empty(sf[-1], np.int64)
    for s in range(len(sf) - 1):
        snode_of[sf[s]:sf[s + 1]] = s
    for lev in range(len(plan["level_ptr"]) - 1):
        for s in plan["level_nodes"][plan["level_ptr"][lev]:plan["level_ptr"][lev + 1]]:
            w = sf[s + 1] - sf[s]
            R = rows[rptr[s]:rptr[s + 1]]
            nr = R.size
            P = panels[poff[s]:poff[s + 1]].reshape(w, nr).T  # view, column-major
            D = np.tril(P[:w, :w])
            D = D + np.tril(D, -1).T
            try:
                Lss = np.linalg.cholesky(D)
            except np.linalg.LinAlgError:
                return sf[s]
            P[:w, :w] = Lss
            if nr > w:
                B = P[w:, :]
                B[:] = np.linalg.solve(Lss, B.T).T
                U = B @ B.T
                Rb = R[w:]
                for b in range(nr - w):
                    col = Rb[b]
                    t = snode_of[col]
                    c = col - sf[t]
                    Rt = rows[rptr[t]:rptr[t + 1]]
                    pos = np.searchsorted(Rt, Rb[b:])
                    assert np.all(Rt[pos] == Rb[b:])
                    base = poff[t] + c * Rt.size
                    panels[base + pos] -= U[b:, b]
    return -1


def panels_to_dense_L(plan, panels):
    m = plan["sfirst"][-1]
    L = np.zeros((m, m))
    sf, rptr, rows, poff = plan["sfirst"], plan["rptr"], plan["rows"], plan["poff"]
    for s in range(len(sf) - 1):
        w = sf[s + 1] - sf[s]
        R = rows[rptr[s]:rptr[s + 1]]
        P = panels[poff[s]:poff[s + 1]].reshape(w, R.size).T
        for c in range(w):
            L[R[c:], sf[s] + c] = P[c:, c]
    return L
