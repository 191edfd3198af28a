"""Verification on the device against AnalyticalSolutions/Diffusion.h, and the standalone
Krylov primitives.

Diffusion_N0D1 (AnalyticalSolutions/Diffusion.h:46-58): a slab with zero gradient at zeta = 0
and phi = 1 at zeta = 1, started from phi = 0, solved step by step through the device path
(dare_csr_fill_values -> dare_bicgstab2_solve -> dare_copy_to_field): along a 1-D grid (CSR
traversal) and along the x axis of a 3-D grid with Neumann walls elsewhere (diagonal-view
traversal). Every step agrees with the oracle's step to 1e-12; the final field matches the
series to the discretisation error (O(dt) + O(dx^2), 3e-3 here).

dare_dot and dare_axpby (the Tpetra dot and update calls of BiCGStab2.inl) are checked
against the oracle's blocked dot and numpy's elementwise update.
"""
import math

import numpy as np
import pytest

import dare_amd as da
import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [[40], [40, 3, 4]])
def test_diffusion_n0d1_slab_on_device(ctx, shape):
    import torch
    n, dt, steps = shape[0], 1e-4, 1000
    dim = len(shape)
    bc = [2] * (2 * dim)
    bc[1] = 1                                    # E face: Dirichlet 1; W and the rest: Neumann
    bv = [0.0] * (2 * dim)
    bv[1] = 1.0
    go = orc.grid(shape, ghost=2)
    g = da.Grid.make(shape, ghost=2)
    to = orc.terms(dt=dt, diff=1.0, bc_type=bc, bc_value=bv)
    td = da.Terms.make(dt=dt, diff=1.0, bc_type=bc, bc_value=bv)
    npad = g.num_padded
    phi = torch.zeros(npad, dtype=torch.float64, device="cuda")
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, td)
    rows = g.num_cells
    b = torch.zeros(rows, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    worst = 0.0
    for s in range(steps):
        prev = phi.cpu().numpy()
        da.fill_values(ctx, A, g, td, phi, b, x)
        r = da.bicgstab2(ctx, A, b, x, tol=1e-13, tol_kind=0, max_it=500)
        assert r["converged"] == 1, s
        da.copy_to_field(ctx, g, x, phi)
        # the oracle's step from the same field
        rp, col, val, rhs, x0 = orc.assemble(go, to, prev)
        xo, ok, *_ = orc.bicgstab2(rp, col, val, rhs, x0, 1e-13, 500)
        assert ok
        exp = orc.copy_to_field(go, xo, prev)
        worst = max(worst, float(np.max(np.abs(phi.cpu().numpy() - exp))))
    assert worst <= 1e-12, worst
    zeta = (np.arange(n) + 0.5) / n
    exact = np.array([orc.lib().orc_diffusion_n0d1(steps * dt, z, 100) for z in zeta])
    field = phi.cpu().numpy().reshape([m + 4 for m in shape])
    inner = field[(slice(2, -2),) * dim]
    prof = inner.reshape(n, -1)
    assert np.max(np.abs(prof - prof[:, :1])) <= 1e-12     # the slab stays one-dimensional
    err = float(np.max(np.abs(prof[:, 0] - exact)))
    assert err < 3e-3, err
    A.close()


def test_dot_and_axpby(ctx, rng):
    import torch
    for n in (1, 255, 256 * 4 * 3 + 7, 1_000_003, 5_000_000):
        a = rng.normal(size=n)
        bvec = rng.normal(size=n)
        got = da.dot(ctx, torch.from_numpy(a).cuda(), torch.from_numpy(bvec).cuda())
        exact = math.fsum(a * bvec)
        scale = float(np.abs(a) @ np.abs(bvec))
        assert abs(got - exact) <= 1e-14 * scale, n
        assert abs(got - orc.lib().orc_dot(n, a.ctypes.data, bvec.ctypes.data)) <= 1e-14 * scale
        y = torch.from_numpy(bvec.copy()).cuda()
        da.check(da.lib().dare_axpby(ctx.h, n, 0.75, torch.from_numpy(a).cuda().data_ptr(), -1.25, y.data_ptr()),
                 "dare_axpby")
        exp = 0.75 * a + -1.25 * bvec
        assert np.array_equal(y.cpu().numpy().view(np.int64), exp.view(np.int64)), n
    with pytest.raises(da.DareError):
        da.check(da.lib().dare_axpby(ctx.h, -1, 1.0, 1, 1.0, 1), "dare_axpby")


def test_stream_probe(ctx, rng):
    """The bandwidth probe computes what it claims (copy, and sums of 8 strided chunks)."""
    import torch
    n = 512 * 300
    src = torch.from_numpy(rng.normal(size=8 * n)).cuda()
    dst = torch.empty(n, dtype=torch.float64, device="cuda")
    da.stream_probe(ctx, src[:n], dst, ratio=1)
    assert torch.equal(dst, src[:n])
    s = src.cpu().numpy().reshape(-1, 8, 256, 2)          # [block][u][lane][2]
    exp = s[:, 0].copy()
    for u in range(1, 8):
        exp = exp + s[:, u]
    for layout in (0, 1):
        dst.zero_()
        da.stream_probe(ctx, src, dst, ratio=8, layout=layout)
        assert np.array_equal(dst.cpu().numpy(), exp.reshape(-1)), layout
    a = src.cpu().numpy().reshape(8, n)                    # layout 2: eight separate arrays
    exp = a[0].copy()
    for u in range(1, 8):
        exp = exp + a[u]
    da.stream_probe(ctx, src, dst, ratio=8, layout=2, grid=64)
    assert np.array_equal(dst.cpu().numpy(), exp)
    # layout 3 (read-only): lane t of work-group b writes the total of its chunks at 256 b + t
    grid = 64
    dst.fill_(-1.0)
    da.stream_probe(ctx, src, dst, ratio=8, layout=3, grid=grid)
    e2 = exp.reshape(-1, 256, 2)                           # [chunk][lane][2]
    nblk = e2.shape[0]
    tot = np.zeros((grid, 256, 2))
    for b in range(grid):
        x = b % 8
        for blk in range(nblk * x // 8 + b // 8, nblk * (x + 1) // 8, grid // 8):
            tot[b] += e2[blk]
    got = dst.cpu().numpy()
    assert np.allclose(got[:grid * 512], tot.reshape(-1), rtol=1e-12, atol=0)
    assert np.all(got[grid * 512:] == -1.0)
