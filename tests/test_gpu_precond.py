"""Preconditioners. Jacobi right preconditioning (precond = 1): the solver runs on A M^-1, M = diag(A), with the
column-scaled diagonal view (or scaled CSR values for a generic matrix), and returns x = M^-1 y. The solution is the same x* (checked
against the oracle's unpreconditioned solve at a tight tolerance), the stopping test is on the
true residual, and on a variable-diffusivity operator it needs fewer passes. Left Jacobi
(precond = 3) runs on diag(A)^-1 A x = diag(A)^-1 b; the multigrid V-cycle (precond = 2) in the
right slot."""
import numpy as np
import pytest

import dare_amd as da
import oracle as orc
from test_gpu_parity import DIA_CASES, device_assembly, dev, host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", DIA_CASES)
def test_jacobi_fp64_matches_oracle(ctx, rng, name):
    go, A, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    tol = 1e-12 * np.linalg.norm(orhs)
    xo, ok, it, res, _ = orc.bicgstab2(rp, col, val, orhs, ox0, tol, 300)
    assert ok
    r = da.bicgstab2(ctx, A, rhs, x0, tol=tol, tol_kind=0, max_it=300, precond=1)
    assert r["converged"] == 1
    x = host(x0)
    assert np.max(np.abs(x - xo)) <= 1e-10 * np.max(np.abs(xo))
    assert np.linalg.norm(orhs - orc.spmv(rp, col, val, x)) <= 1.001 * tol
    A.close()


@pytest.mark.parametrize("name", DIA_CASES)
def test_jacobi_mixed_reaches_fp64_residual(ctx, rng, name):
    go, A, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    r = da.refine_mixed(ctx, A, rhs, x0, tol=1e-10, tol_kind=1, inner_tol=1e-4, precond=1)
    assert r["converged"] == 1
    x = host(x0)
    assert np.linalg.norm(orhs - orc.spmv(rp, col, val, x)) <= 1e-10 * np.linalg.norm(orhs)
    A.close()


def test_jacobi_saves_passes_on_variable_diffusivity(ctx):
    import torch
    from dare_amd.workloads import c3_lognormal
    n = 48
    D, src, dt = c3_lognormal(n, seed=11, mean_diff_number=4.0)
    g = da.Grid.make([n] * 3, size=[1, 1, 1], ghost=2)
    t = da.Terms.make(dt=dt, diff_cell=D, source=src, bc_type=[1] * 6, bc_value=[0.0] * 6)
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, t)
    phi = torch.zeros((n + 4) ** 3, dtype=torch.float64, device="cuda")
    rhs = torch.zeros(n ** 3, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(rhs)
    da.fill_values(ctx, A, g, t, phi, rhs, x)
    r0 = da.bicgstab2(ctx, A, rhs, x.clone(), tol=1e-10, tol_kind=1, max_it=1000)
    xj = x.clone()
    r1 = da.bicgstab2(ctx, A, rhs, xj, tol=1e-10, tol_kind=1, max_it=1000, precond=1)
    assert r0["converged"] == 1 and r1["converged"] == 1
    assert r1["iterations"] < r0["iterations"], (r0["iterations"], r1["iterations"])
    y = torch.empty_like(xj)
    da.spmv(ctx, A, xj, y)
    assert float(torch.linalg.norm(rhs - y) / torch.linalg.norm(rhs)) <= 1e-10
    A.close()


@pytest.mark.parametrize("name", ["periodic_even", "wide_dirichlet", "react2"])
def test_jacobi_on_generic_csr_matches_oracle(ctx, rng, name):
    """A matrix uploaded as plain CSR (the generic host-functor path, no diagonal view): Jacobi
    finds the diagonal in the rows and scales a copy of the values; FP64 and mixed solves reach
    the oracle's solution / the FP64 tolerance."""
    go, Ad, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    Ad.close()
    A = da.Csr(ctx)
    da.upload_host_csr(ctx, A, rp, col, val)
    tol = 1e-12 * np.linalg.norm(orhs)
    xo, ok, it, res, _ = orc.bicgstab2(rp, col, val, orhs, ox0, tol, 300)
    assert ok
    x = dev(ox0)
    r = da.bicgstab2(ctx, A, dev(orhs), x, tol=tol, tol_kind=0, max_it=300, precond=1)
    assert r["converged"] == 1
    xh = host(x)
    assert np.max(np.abs(xh - xo)) <= 1e-10 * np.max(np.abs(xo))
    assert np.linalg.norm(orhs - orc.spmv(rp, col, val, xh)) <= 1.001 * tol
    xm = dev(ox0)
    r = da.refine_mixed(ctx, A, dev(orhs), xm, tol=1e-10, tol_kind=1, inner_tol=1e-4, precond=1)
    assert r["converged"] == 1
    assert np.linalg.norm(orhs - orc.spmv(rp, col, val, host(xm))) <= 1e-10 * np.linalg.norm(orhs)
    A.close()


def _left_scaled(rp, col, val, b):
    """diag(A)^-1 A and diag(A)^-1 b, row by row (the system left Jacobi iterates on)."""
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    diag = np.zeros(n)
    on = col == rows
    diag[rows[on]] = val[on]
    dinv = 1.0 / diag
    return val * dinv[rows], b * dinv


@pytest.mark.parametrize("name", DIA_CASES)
def test_left_jacobi_is_bicgstab2_on_the_row_scaled_system(ctx, rng, name):
    """precond = 3 (Jacobi in the left slot, BiCGStab2.inl:276-280): the oracle's BiCGStab2 on
    (diag(A)^-1 A, diag(A)^-1 b) takes the same passes (+-1) to the same solution, and that
    solution solves A x = b."""
    go, A, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    vs, bs = _left_scaled(rp, col, val, orhs)
    tol = 1e-12 * np.linalg.norm(bs)
    xo, ok, it, res, _ = orc.bicgstab2(rp, col, vs, bs, ox0, tol, 300)
    assert ok
    r = da.bicgstab2(ctx, A, rhs, x0, tol=tol, tol_kind=0, max_it=300, precond=da.PRECOND_JACOBI_LEFT)
    assert r["converged"] == 1 and abs(r["iterations"] - it) <= 1, (r["iterations"], it)
    assert r["residual"] <= tol
    x = host(x0)
    assert np.max(np.abs(x - xo)) <= 1e-10 * np.max(np.abs(xo))
    assert np.linalg.norm(orhs - orc.spmv(rp, col, val, x)) <= 1e-9 * np.linalg.norm(orhs)
    A.close()


def test_left_jacobi_on_generic_csr_and_refusals(ctx, rng):
    """Left Jacobi on a plain CSR upload (rows scaled in a copy of the values); the mixed
    solve and the allocation-free entry point refuse it."""
    go, Ad, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, "wide_dirichlet")
    Ad.close()
    A = da.Csr(ctx)
    da.upload_host_csr(ctx, A, rp, col, val)
    vs, bs = _left_scaled(rp, col, val, orhs)
    tol = 1e-12 * np.linalg.norm(bs)
    xo, ok, it, res, _ = orc.bicgstab2(rp, col, vs, bs, ox0, tol, 300)
    x = dev(ox0)
    r = da.bicgstab2(ctx, A, dev(orhs), x, tol=tol, tol_kind=0, max_it=300, precond=da.PRECOND_JACOBI_LEFT)
    assert r["converged"] == 1 and abs(r["iterations"] - it) <= 1
    assert np.max(np.abs(host(x) - xo)) <= 1e-10 * np.max(np.abs(xo))
    with pytest.raises(da.DareError, match="precond must be"):
        da.refine_mixed(ctx, A, dev(orhs), dev(ox0), tol=1e-10, tol_kind=1, precond=da.PRECOND_JACOBI_LEFT)
    A.close()


def test_precision_form_error_vs_fp64_reference(ctx):
    """FP32 inner solves inside FP64 refinement to true relative residual 1e-10, variable
    diffusivity at 64^3: the error against an FP64 solve to 1e-13 stays at the level the
    residual allows, with and without Jacobi."""
    import torch
    from dare_amd.workloads import c3_lognormal
    n = 64
    D, src, dt = c3_lognormal(n, seed=3, mean_diff_number=1.0)
    g = da.Grid.make([n] * 3, size=[1, 1, 1], ghost=2)
    t = da.Terms.make(dt=dt, diff_cell=D, source=src, bc_type=[1] * 6, bc_value=[0.0] * 6)
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, t)
    phi = torch.zeros((n + 4) ** 3, dtype=torch.float64, device="cuda")
    rhs = torch.zeros(n ** 3, dtype=torch.float64, device="cuda")
    x0 = torch.zeros_like(rhs)
    da.fill_values(ctx, A, g, t, phi, rhs, x0)
    xr = x0.clone()
    assert da.bicgstab2(ctx, A, rhs, xr, tol=1e-13, tol_kind=1, max_it=5000)["converged"] == 1
    for pre in (0, 1):
        x = x0.clone()
        r = da.refine_mixed(ctx, A, rhs, x, tol=1e-10, tol_kind=1, inner_tol=1e-4, max_it=10000, precond=pre)
        assert r["converged"] == 1 and r["refinements"] >= 2
        err = float(torch.linalg.norm(x - xr) / torch.linalg.norm(xr))
        print(f"precond {pre}: passes {r['iterations']} refinements {r['refinements']} error {err:.2e}")
        assert err <= 1e-8
    A.close()


# ---- multigrid (precond = 2): aggregation V-cycle on the diagonal view, the MueLu "AMG" stand-in
MG_CASES = [c for c in DIA_CASES if c not in ("react2", "fixed_react2")]


@pytest.mark.parametrize("name", MG_CASES)
def test_multigrid_fp64_matches_oracle(ctx, rng, name):
    """Right-preconditioned FP64 BiCGStab2 with the V-cycle: the oracle's unpreconditioned
    solution at a tight tolerance, and the true residual below it."""
    go, A, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    tol = 1e-12 * np.linalg.norm(orhs)
    xo, ok, it, res, _ = orc.bicgstab2(rp, col, val, orhs, ox0, tol, 300)
    assert ok
    r = da.bicgstab2(ctx, A, rhs, x0, tol=tol, tol_kind=0, max_it=300, precond=da.PRECOND_MG)
    assert r["converged"] == 1, r
    x = host(x0)
    assert np.max(np.abs(x - xo)) <= 1e-10 * np.max(np.abs(xo))
    assert np.linalg.norm(orhs - orc.spmv(rp, col, val, x)) <= 1.001 * tol
    A.close()


@pytest.mark.parametrize("name", MG_CASES)
def test_multigrid_mixed_reaches_fp64_residual(ctx, rng, name):
    go, A, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    r = da.refine_mixed(ctx, A, rhs, x0, tol=1e-10, tol_kind=1, inner_tol=1e-4, precond=da.PRECOND_MG)
    assert r["converged"] == 1, r
    x = host(x0)
    assert np.linalg.norm(orhs - orc.spmv(rp, col, val, x)) <= 1e-10 * np.linalg.norm(orhs)
    A.close()


def test_multigrid_refuses_two_components(ctx, rng):
    go, A, rhs, x0, ref, keep = device_assembly(ctx, rng, "react2")
    with pytest.raises(da.DareError, match="one-component operator"):
        da.bicgstab2(ctx, A, rhs, x0, tol=1e-10, tol_kind=1, precond=da.PRECOND_MG)
    A.close()


def test_multigrid_saves_passes_on_variable_diffusivity(ctx):
    """log-normal diffusivity (sigma 1), mean diffusion number 4: the V-cycle needs fewer
    passes than Jacobi, which needs fewer than no preconditioner; all reach 1e-10."""
    import torch
    from dare_amd.workloads import c3_lognormal
    n = 48
    D, src, dt = c3_lognormal(n, seed=11, mean_diff_number=4.0)
    g = da.Grid.make([n] * 3, size=[1, 1, 1], ghost=2)
    t = da.Terms.make(dt=dt, diff_cell=D, source=src, bc_type=[1] * 6, bc_value=[0.0] * 6)
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, t)
    phi = torch.zeros((n + 4) ** 3, dtype=torch.float64, device="cuda")
    rhs = torch.zeros(n ** 3, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(rhs)
    da.fill_values(ctx, A, g, t, phi, rhs, x)
    its = {}
    for pc in (0, 1, 2):
        xs = x.clone()
        r = da.bicgstab2(ctx, A, rhs, xs, tol=1e-10, tol_kind=1, max_it=1000, precond=pc)
        assert r["converged"] == 1, (pc, r)
        y = torch.empty_like(xs)
        da.spmv(ctx, A, xs, y)
        assert float(torch.linalg.norm(rhs - y) / torch.linalg.norm(rhs)) <= 1e-10
        its[pc] = r["iterations"]
    print("passes (none, Jacobi, multigrid):", its)
    assert its[2] < its[1] < its[0], its
    A.close()
