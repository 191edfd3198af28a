"""The allocation-free FP64 solve (dare_bicgstab2_solve_noalloc) against dare_bicgstab2_solve.

Both run the same kernels; the allocation-free one keeps its work vectors, reduction partials
and scalars in the library's static device scratch. Its results must be bit-identical to the
existing entry point's on two calls in a row (the second on a smaller system, so the scratch
still holds the first call's data), it must not change the device's free memory, and it must
refuse systems beyond its maximum and preconditioned solves from the host.
"""
import ctypes

import numpy as np
import pytest

import dare_amd as da

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 18


def _system(ctx, n, periodic=None, conv=False, seed=0):
    import torch
    rng = np.random.default_rng(seed)
    g = da.Grid.make(n, ghost=2, periodic=periodic)
    phi = torch.from_numpy(rng.uniform(size=g.num_padded)).cuda()
    kw = dict(dt=1e-2, diff=1e-3)
    if conv:
        vel = [torch.from_numpy(rng.uniform(-1, 1, size=g.num_faces(d))).cuda() for d in range(g.dim)]
        kw.update(limiter=da.VANLEER, vel=vel)
    if periodic is None or not all(periodic):
        kw.update(bc_type=[0 if (periodic and periodic[f // 2]) else da.BC_DIRICHLET for f in range(2 * g.dim)],
                  bc_value=[0.25 * f for f in range(2 * g.dim)])
    t = da.Terms.make(**kw)
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, t)
    rows = g.num_cells
    rhs = torch.zeros(rows, dtype=torch.float64, device="cuda")
    x0 = torch.zeros_like(rhs)
    da.fill_values(ctx, A, g, t, phi, rhs, x0)
    return A, rhs, x0, t


def _both(ctx, A, rhs, x0, tol):
    import torch
    xa = x0.clone()
    ra = da.bicgstab2(ctx, A, rhs, xa, tol=tol, tol_kind=1, max_it=200, history=64)
    xb = x0.clone()
    free0 = torch.cuda.mem_get_info()[0]
    rb = da.bicgstab2_noalloc(ctx, A, rhs, xb, tol=tol, tol_kind=1, max_it=200, history=64)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    return xa.cpu().numpy(), ra, xb.cpu().numpy(), rb, free0, free1


def test_noalloc_solve_matches_the_allocating_solve_twice(ctx):
    # call 1: the largest allowed system, 64^3 (diagonal-view operator)
    print("noalloc call 1: 64^3 diffusion, 262144 rows", flush=True)
    A, rhs, x0, _ = _system(ctx, [64, 64, 64], seed=1)
    assert A.nrows == MAX_ROWS
    xa, ra, xb, rb, f0, f1 = _both(ctx, A, rhs, x0, 1e-10)
    assert ra["converged"] == 1 and rb["converged"] == 1
    assert (ra["iterations"], ra["status"], ra["history"]) == (rb["iterations"], rb["status"], rb["history"])
    assert ra["residual"] == rb["residual"] and ra["bnorm"] == rb["bnorm"]
    assert np.array_equal(xa.view(np.int64), xb.view(np.int64))
    assert f0 == f1, (f0, f1)
    A.close()
    # call 2: a smaller, non-symmetric system (implicit van Leer convection, one periodic axis)
    print("noalloc call 2: [20,18,16] convection-diffusion", flush=True)
    B, rhs, x0, _ = _system(ctx, [20, 18, 16], periodic=[1, 0, 0], conv=True, seed=2)
    xa, ra, xb, rb, f0, f1 = _both(ctx, B, rhs, x0, 1e-11)
    assert ra["converged"] == 1 and rb["converged"] == 1
    assert (ra["iterations"], ra["history"]) == (rb["iterations"], rb["history"])
    assert np.array_equal(xa.view(np.int64), xb.view(np.int64))
    assert f0 == f1, (f0, f1)
    B.close()


def test_noalloc_solve_generic_csr_and_refusals(ctx):
    # a 2-D system has no diagonal view: the CSR traversal runs from the scratch as well
    print("noalloc: 2-D CSR system [37,29]", flush=True)
    C, rhs, x0, _ = _system(ctx, [37, 29], seed=3)
    xa, ra, xb, rb, f0, f1 = _both(ctx, C, rhs, x0, 1e-12)
    assert ra["converged"] == 1 and (ra["iterations"], ra["history"]) == (rb["iterations"], rb["history"])
    assert np.array_equal(xa.view(np.int64), xb.view(np.int64))
    assert f0 == f1
    # refused on the host, before any kernel: preconditioning, and more rows than the scratch holds
    import torch
    p, _ = da._params(1e-8, 0, 10, 0.0, 0, 0, 0, 1)
    r = da.SolverResult()
    xp = x0.clone()
    rc = da.lib().dare_bicgstab2_solve_noalloc(ctx.h, C.h, da._ptr(rhs), da._ptr(xp), ctypes.byref(p),
                                               ctypes.byref(r))
    assert rc == -6 and b"precond must be 0" in da.lib().dare_last_error()
    C.close()
    D, rhs, x0, _ = _system(ctx, [65, 64, 64], seed=4)
    assert D.nrows > MAX_ROWS
    with pytest.raises(da.DareError, match="allocation-free maximum"):
        da.bicgstab2_noalloc(ctx, D, rhs, x0.clone(), tol=1e-8)
    D.close()
    torch.cuda.synchronize()
