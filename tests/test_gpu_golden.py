"""The HIP path against the committed golden fixtures (tests/golden/).

Operator H v to 1e-13 (reduction order), P^-1 v with the fixture's core to
1e-10 (LAPACK solve of the same matrix), the selector network's FP32
gradient to 1e-4 of its norm and BF16 logits to 5e-2 (1 + |z|), and the LP
optima (HiGHS) to 1e-8 for the FP64, mixed and mixed + block paths.
"""
import ctypes as C
import os

import numpy as np
import pytest

import mf_psipm_amd as P
from mf_psipm_amd.preconditioner import HipSpectralStableColumnPreconditioner
from mf_psipm_amd.selector import LearnedColumnSelector
from mf_psipm_amd.symbolic import SymbolicPlan

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_operator_and_preconditioner_match_golden(gpu, precision):
    _say(f"golden operator 400x1200 {precision}")
    g = _load("operator_400x1200.npz")
    raw, *_ = P.generators.random_lp(400, 1200, 5, 11)
    lp = P.lp.to_standard_form(raw)
    P.lp.scale_problem(lp)
    obj = HipSpectralStableColumnPreconditioner(lp.A, protected=~lp.pos, precision=precision)
    obj._call("psg_set_weights", obj._ctx, g["op_theta"], float(g["op_delta"]))
    Hv = obj._full_matvec(g["op_v"])
    assert np.allclose(Hv, g["op_Hv"], rtol=1e-13, atol=1e-13 * np.abs(g["op_Hv"]).max())
    plan = SymbolicPlan(lp.A, g["op_core"])
    assert obj._call("psg_factor", obj._ctx, plan.handle, 0.0, None) == 0
    z = obj._apply(g["op_v"])
    tol = 1e-10 if precision == "fp64" else 1e-4
    assert np.linalg.norm(z - g["op_Pinv_v"]) <= tol * np.linalg.norm(g["op_Pinv_v"])
    plan.close()
    obj.close()


def test_selector_network_matches_golden(gpu):
    _say("golden selector network")
    g = _load("mlp_selector.npz")
    norm = np.concatenate([np.zeros(8), np.ones(8)]).astype(np.float32)
    net = LearnedColumnSelector(g["mlp_params"], norm)
    loss, grad = net.loss_grad(g["mlp_X"], g["mlp_y"], 2.0)
    assert abs(loss - float(g["mlp_loss"])) <= 1e-5 * (1.0 + abs(float(g["mlp_loss"])))
    assert np.linalg.norm(grad - g["mlp_grad"]) <= 1e-4 * np.linalg.norm(g["mlp_grad"])
    z = net.forward(g["mlp_X"])
    assert np.all(np.abs(z - g["mlp_logits"]) <= 5e-2 * (1.0 + np.abs(g["mlp_logits"])))
    net.close()


@pytest.mark.parametrize("seed,precision,pc", [(0, "fp64", 3), (1, "fp64", 3), (1, "mixed", 4)])
def test_lp_optimum_matches_golden(gpu, seed, precision, pc):
    _say(f"golden LP 2000x6000 seed {seed} {precision} pc={pc}")
    g = _load("lp_optima.npz")
    raw, *_ = P.generators.random_lp(2000, 6000, 5, seed)
    lp = P.lp.to_standard_form(raw)
    P.lp.scale_problem(lp)
    tol = P.lp.scaled_tolerance(lp, 1e-8)
    solver = HipSpectralStableColumnPreconditioner(lp.A, protected=~lp.pos, precision=precision,
                                                   cg_tol=1e-8, cg_tol_max=1e-8)
    sol = P.ipm.psipm_solve(lp, solver, tol=tol, compl="gap", maxit=200, pc=pc)
    info = P.lp.recover(lp, raw, sol.x, sol.y, sol.z, sol.w)
    ref = float(g[f"random_2000x6000_s{seed}_obj"])
    assert sol.status == "optimal"
    assert abs(info["obj"] - ref) <= 1e-8 * abs(ref)
    assert max(info["pinf"], info["dinf"]) <= 1e-8
    solver.close()


def test_setcover_optimum_matches_golden(gpu):
    _say("golden set cover 300x6000")
    from mf_psipm_amd.selector import solve_with
    g = _load("lp_optima.npz")
    raw = P.generators.set_covering(300, 6000, 7)
    lp = P.lp.to_standard_form(raw)
    P.lp.scale_problem(lp)
    tol = P.lp.scaled_tolerance(lp, 1e-8)
    sol, st = solve_with(lp, tol, precision="mixed")
    info = P.lp.recover(lp, raw, sol.x, sol.y, sol.z, sol.w)
    ref = float(g["setcover_300x6000_s7_obj"])
    assert st["status"] == "optimal"
    assert abs(info["obj"] - ref) <= 1e-8 * abs(ref)
