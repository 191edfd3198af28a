"""The HIP path against the committed golden fixtures (tests/golden/), with no oracle build:
stencils, ghost fill, face interpolation, TVD and the CSR assembly bit-exact; BiCGStab2 within
one pass of the fixture's count, the solution within 1e-12 relative and the residual norms
of passes 1-3 within 1e-10 relative or 1e-13 of ||b|| (the dot products are summed in another
order)."""
import os
import sys

import numpy as np
import pytest

import dare_amd as da
from conftest import bits_equal

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from spec import GHOST_MODES, GHOST_VALUES, LIMITERS, operators  # noqa: E402

pytestmark = pytest.mark.gpu


def load(n):
    with np.load(os.path.join(HERE, "golden", f"golden_{n}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def zeros_like(a):
    import torch
    return torch.zeros(a.size, dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("n", [8, 16])
def test_stencils_match_golden(ctx, n):
    Z = load(n)
    g = da.Grid.make([n, n, n], ghost=2)
    phi = dev(Z["phi"])
    vel = [dev(Z[f"vel_{a}"]) for a in "xyz"]
    f = [zeros_like(Z[f"grad_{a}"]) for a in "xyz"]
    da.grad_faces(ctx, g, phi, f)
    for d, a in enumerate("xyz"):
        assert bits_equal(f[d].cpu().numpy(), Z[f"grad_{a}"]), a
    out = zeros_like(Z["div"])
    da.div_faces(ctx, g, f, out)
    assert bits_equal(out.cpu().numpy(), Z["div"])
    for name, lid in LIMITERS.items():
        if f"tvd_{name}_x" not in Z:
            continue
        da.tvd_face_flux(ctx, g, lid, phi, vel, f)
        for d, a in enumerate("xyz"):
            assert bits_equal(f[d].cpu().numpy(), Z[f"tvd_{name}_{a}"]), (name, a)
    step = phi.clone()
    da.tvd_explicit_step(ctx, g, da.VANLEER, phi, vel, 1e-3, step)
    assert bits_equal(step.cpu().numpy(), Z["step_vanleer"])
    gh = phi.clone()
    da.ghost_fill(ctx, g, gh, GHOST_MODES, GHOST_VALUES)
    assert bits_equal(gh.cpu().numpy(), Z["ghost"])
    if "ucell_x" in Z:
        da.interpolate_to_faces(ctx, g, [dev(Z[f"ucell_{a}"]) for a in "xyz"], f)
        for d, a in enumerate("xyz"):
            assert bits_equal(f[d].cpu().numpy(), Z[f"uface_{a}"]), a


@pytest.mark.parametrize("n", [8, 16])
def test_assembly_and_solve_match_golden(ctx, n):
    import torch
    Z = load(n)
    vel = [dev(Z[f"vel_{a}"]) for a in "xyz"]
    for name, ncomp, kw, extra in operators(n):
        g = da.Grid.make([n, n, n], ghost=2, ncomp=ncomp)
        arrays = {}
        if "diff_cell" in extra:
            arrays["diff_cell"] = dev(Z[f"{name}_diffcell"])
        if "source" in extra:
            arrays["source"] = dev(Z[f"{name}_source"])
        if "vel" in extra:
            arrays["vel"] = vel
        if "react" in extra:
            arrays["react"] = dev(Z[f"{name}_react"])
        t = da.Terms.make(**kw, **arrays)
        field = dev(Z["phi"] if ncomp == 1 else Z[f"{name}_phi"])
        A = da.Csr(ctx)
        da.build_structure(ctx, A, g, t)
        rows = len(Z[f"{name}_rhs"])
        b = torch.zeros(rows, dtype=torch.float64, device="cuda")
        x = torch.zeros_like(b)
        da.fill_values(ctx, A, g, t, field, b, x)
        rp, col, val = A.download()
        assert np.array_equal(rp, Z[f"{name}_rp"]) and np.array_equal(col, Z[f"{name}_col"]), name
        assert bits_equal(val, Z[f"{name}_val"]), name
        assert bits_equal(b.cpu().numpy(), Z[f"{name}_rhs"]) and bits_equal(x.cpu().numpy(), Z[f"{name}_x0"]), name
        r = da.bicgstab2(ctx, A, b, x, tol=float(Z[f"{name}_tol"][0]), tol_kind=0, max_it=500, history=3)
        assert r["converged"] == 1 and abs(r["iterations"] - int(Z[f"{name}_it"][0])) <= 1, (name, r["iterations"])
        xs = Z[f"{name}_x"]
        assert np.max(np.abs(x.cpu().numpy() - xs)) <= 1e-12 * np.max(np.abs(xs)), name
        # residual norms of passes 1-3, to 1e-10 relative or 1e-13 of ||b|| (a norm far below ||b||
        # is a cancellation result: its relative rounding differences grow as it shrinks)
        h = np.array(r["history"][:len(Z[f"{name}_hist"])])
        href = Z[f"{name}_hist"]
        bnorm = float(np.linalg.norm(Z[f"{name}_rhs"]))
        assert len(h) == len(href) and np.all(np.abs(h - href) <= 1e-10 * href + 1e-13 * bnorm), (name, h, href)
        A.close()
