"""GPU test of the reduced HSD Newton system on the spectral KKT object:
the direction from the block (f, h) solve plus the theta-row solve equals
the oracle's exact-solve direction, the corrector reuses the cached rows,
and three single solves give the same direction."""
import os
import sys

import numpy as np
import pytest

from mf_psipm_amd.hsd import HSDEmbedding, HSDNewtonSystem, HSDState
from mf_psipm_amd.preconditioner import HipSpectralStableColumnPreconditioner
from oracle import hsd_cpu as O  # noqa: E402

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(__file__))
from test_hsd_cpu import _problem  # noqa: E402


@pytest.mark.parametrize("pair", [True, False])
def test_hsd_direction_matches_oracle(gpu, pair):
    print(f"\n[gpu-case] HSD direction 1200x3600 pair={pair}", flush=True)
    A, c, upper, b, D, st, emb, rhs = _problem(1200, 3600, 4)
    kkt = HipSpectralStableColumnPreconditioner(A, structural_backbone=False, cg_maxit=3000)
    sys_ = HSDNewtonSystem(kkt, A, c, upper, b, D, 1e-6, 1e-4, HSDState(st.x, st.z, st.w, st.tau, st.kappa),
                           HSDEmbedding(emb.bar_b, emb.bar_c, emb.bar_g), pcg_tolerance=1e-11, pair=pair)
    got = sys_.direction(*rhs)
    ref = O.direction(A, c, upper, b, D, 1e-6, 1e-4, st, emb, 1e-8, rhs)
    assert got[-1]
    for g, r in zip(got[:4], ref[:4]):
        assert np.linalg.norm(g - r) <= 1e-7 * max(1.0, np.linalg.norm(r))
    for g, r in zip(got[4:7], ref[4:7]):
        assert abs(g - r) <= 1e-7 * max(1.0, abs(r))
    assert sys_.solves == 3 and sys_.pair_solves == (1 if pair else 0)
    rhs2 = tuple(v * 0.5 if isinstance(v, float) else v[::-1].copy() for v in rhs)
    got2 = sys_.direction(*rhs2)
    ref2 = O.direction(A, c, upper, b, D, 1e-6, 1e-4, st, emb, 1e-8, rhs2)
    assert sys_.solves == 4
    assert np.linalg.norm(got2[1] - ref2[1]) <= 1e-7 * np.linalg.norm(ref2[1])
    kkt.close()
