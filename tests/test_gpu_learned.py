"""What the trained warm-start network gains (weights trained by dare_amd.warmstart on
plain-solver trajectories, committed under dare_amd/weights/).

On held-out trajectories the solver started from the network's guess needs fewer BiCGStab2
passes than from the reference's guess phi^n, and every guess still reaches the same FP64
true relative residual: the plain method stays the judge."""
import os

import numpy as np
import pytest

import dare_amd as da
from dare_amd.learned import WarmStartNet
from dare_amd.warmstart import evaluate

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dare_amd", "weights")


@pytest.mark.parametrize("workload,fname,n", [("advdiff", "ws_advdiff_128.npz", 128),
                                              ("vardiff", "ws_vardiff_128.npz", 128)])
def test_learned_guess_saves_passes(ctx, workload, fname, n):
    path = os.path.join(WEIGHTS, fname)
    if not os.path.exists(path):
        pytest.skip(f"{fname} not trained yet")
    w, meta = WarmStartNet.load_params(path)
    net = WarmStartNet(ctx, [n] * 3, in_channels=meta["in_channels"], params=w)
    ev = evaluate(ctx, net, n=n, eval_seeds=[5000], steps=12, workload=workload)
    print({m: round(ev[m]["passes_mean"], 2) for m in ev})
    assert ev["learned"]["passes_mean"] < ev["previous"]["passes_mean"]
    assert ev["learned"]["passes_saved_vs_previous"] > 0.05
    net.close()


def test_learned_guess_reaches_fp64_tolerance(ctx):
    import torch
    from dare_amd.workloads import C5AdvDiff
    path = os.path.join(WEIGHTS, "ws_advdiff_128.npz")
    if not os.path.exists(path):
        pytest.skip("weights not trained yet")
    w, meta = WarmStartNet.load_params(path)
    net = WarmStartNet(ctx, [128] * 3, in_channels=meta["in_channels"], params=w)
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
