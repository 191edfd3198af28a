"""What the trained warm-start networks gain (weights trained by dare_amd.warmstart on
plain-solver trajectories, committed under dare_amd/weights/ with their training and
evaluation reports).

Every guess still reaches the same FP64 true relative residual: the plain method stays the
judge; what changes is the number of passes."""
import os

import numpy as np
import pytest

import dare_amd as da
from dare_amd.learned import WarmStartNet
from dare_amd.warmstart import evaluate

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dare_amd", "weights")


def test_learned_guess_saves_passes(ctx):
    """Advection-diffusion (128^3, held-out seed): the learned correction of the extrapolated
    guess needs fewer passes than phi^n and than the extrapolation itself, and starts from a
    smaller residual."""
    path = os.path.join(WEIGHTS, "ws_advdiff_128.npz")
    if not os.path.exists(path):
        pytest.skip("weights not trained yet")
    net = WarmStartNet.load(ctx, path, [128] * 3)
    ev = evaluate(ctx, net, n=128, eval_seeds=[5000], steps=16, workload="advdiff")
    print({m: (round(ev[m]["passes_mean"], 2), ev[m]["initial_rel_residual_geomean"]) for m in ev})
    assert all(ev[m]["all_converged"] for m in ev)
    assert ev["learned"]["passes_mean"] < ev["previous"]["passes_mean"]
    assert ev["learned"]["passes_mean"] <= ev["extrapolated"]["passes_mean"]
    assert ev["learned"]["initial_rel_residual_geomean"] < ev["extrapolated"]["initial_rel_residual_geomean"]
    net.close()


def test_learned_guess_variable_diffusivity_reaches_tolerance(ctx):
    """Variable diffusivity (128^3): whatever the guess, every solve reaches the FP64 tolerance.
    (Here the learned correction lowers the initial residual but needs more passes than phi^n:
    its error sits in the slow modes of the operator; the evaluation reports it as measured.)"""
    path = os.path.join(WEIGHTS, "ws_vardiff_128.npz")
    if not os.path.exists(path):
        pytest.skip("weights not trained yet")
    net = WarmStartNet.load(ctx, path, [128] * 3)
    ev = evaluate(ctx, net, n=128, eval_seeds=[5001], steps=6, workload="vardiff")
    assert all(ev[m]["all_converged"] for m in ev)
    assert ev["learned"]["initial_rel_residual_geomean"] < ev["previous"]["initial_rel_residual_geomean"]
    net.close()


def test_learned_guess_reaches_fp64_tolerance(ctx):
    import torch
    from dare_amd.workloads import C5AdvDiff
    path = os.path.join(WEIGHTS, "ws_advdiff_128.npz")
    if not os.path.exists(path):
        pytest.skip("weights not trained yet")
    net = WarmStartNet.load(ctx, path, [128] * 3)
    W = C5AdvDiff(ctx, 5001, n=128)
    xg = torch.empty_like(W.x)
    for k in range(4):
        W.assemble()
        net.predict(W.grid, W.phi, W.phi_prev, W.xb, W.r0, W.rhs_scale, xg)
        r = da.bicgstab2(ctx, W.A, W.rhs, xg, tol=1e-8, tol_kind=1, max_it=500)
        assert r["converged"] == 1
        y = torch.empty_like(xg)
        da.spmv(ctx, W.A, xg, y)
        assert float(torch.linalg.norm(W.rhs - y) / torch.linalg.norm(W.rhs)) <= 1e-8
        W.advance(xg)
    net.close()
