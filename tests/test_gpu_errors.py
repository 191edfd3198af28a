"""Error paths of the C-ABI, forced from the host: every invalid argument below is refused
with DARE_ERR_* and a message (dare_last_error) before any kernel runs on it, as the
reference refuses misuse (TEUCHOS_TEST_FOR_EXCEPTION, BiCGStab2.inl:139-160; Terminate in
Trilinos::Build). After all of them the same context still solves a system."""
import numpy as np
import pytest

import dare_amd as da
from dare_amd.learned import WarmStartNet

pytestmark = pytest.mark.gpu


def _raises(fn, text):
    with pytest.raises(da.DareError) as e:
        fn()
    assert text in str(e.value), str(e.value)


def test_invalid_arguments_are_refused_and_the_context_survives(ctx):
    import torch
    g = da.Grid.make([6, 5, 4], ghost=2)
    phi = torch.zeros(g.num_padded, dtype=torch.float64, device="cuda")
    out = torch.zeros_like(phi)
    vel = [torch.zeros(g.num_faces(d), dtype=torch.float64, device="cuda") for d in range(3)]

    # grid descriptors and stencils
    _raises(lambda: da.ghost_fill(ctx, da.Grid.make([6, 5, 4], ghost=2, ncomp=9), phi, [1] * 6),
            "invalid grid descriptor")
    _raises(lambda: da.ghost_fill(ctx, g, phi, [7] * 6), "invalid ghost mode")
    _raises(lambda: da.ghost_fill(ctx, g, phi, [da.GHOST_PERIODIC] * 6), "periodic ghost fill on a non-periodic axis")
    _raises(lambda: da.tvd_explicit_step(ctx, g, 9, phi, vel, 0.1, out), "unknown limiter")
    _raises(lambda: da.tvd_explicit_step(ctx, da.Grid.make([6, 5, 4], ghost=1), da.VANLEER, phi, vel, 0.1, out),
            "TVD needs >= 2 ghost layers")
    _raises(lambda: da.tvd_explicit_step(ctx, g, da.VANLEER, phi, vel, 0.1, phi), "invalid arguments")

    # assembly: boundary types, limiter, unedited faces, reaction shape, 32-bit ordinals
    A = da.Csr(ctx)
    _raises(lambda: da.build_structure(ctx, A, g, da.Terms.make(dt=1.0, diff=1.0, bc_type=[5] * 6)),
            "invalid boundary type")
    _raises(lambda: da.build_structure(ctx, A, g, da.Terms.make(dt=1.0, diff=1.0, bc_type=[1] * 6, limiter=8,
                                                                vel=vel)), "unknown limiter")
    _raises(lambda: da.build_structure(ctx, A, g, da.Terms.make(dt=1.0, diff=1.0, bc_type=[1, 1, 1, 1, 0, 1])),
            "non-periodic boundary face without a boundary edit")
    react = torch.zeros(4 * 120, dtype=torch.float64, device="cuda")
    _raises(lambda: da.build_structure(ctx, A, g, da.Terms.make(dt=1.0, diff=1.0, bc_type=[1] * 6, react=react)),
            "reaction blocks need ncomp == 2")
    _raises(lambda: da.build_structure(ctx, A, da.Grid.make([1400, 1400, 1400], ghost=2),
                                       da.Terms.make(dt=1.0, diff=1.0, bc_type=[1] * 6)),
            "grid too large for 32-bit ordinals")
    # a single-cell non-periodic axis: the edit of the low face leaves the high neighbour, whose
    # column would lie past the last row
    _raises(lambda: da.build_structure(ctx, A, da.Grid.make([1, 6, 12], ghost=2),
                                       da.Terms.make(dt=1.0, diff=1.0, bc_type=[1] * 6)),
            "single-cell non-periodic axis")
    _raises(lambda: da.build_structure(ctx, A, da.Grid.make([7, 1], ghost=1),
                                       da.Terms.make(dt=1.0, diff=1.0, bc_type=[2] * 4)),
            "single-cell non-periodic axis")
    t = da.Terms.make(dt=1.0, diff=1.0, bc_type=[1] * 6, bc_value=[0.0] * 6)
    da.build_structure(ctx, A, g, t)
    rhs = torch.zeros(120, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(rhs)
    _raises(lambda: da.fill_values(ctx, A, g, da.Terms.make(dt=1.0, diff=1.0, bc_type=[2] * 6), phi, rhs, x),
            "boundary edits differ from the structure's")
    # same cell count, other shape or periodicity: refused (the diagonal view walks the build grid)
    for other in (da.Grid.make([4, 5, 6], ghost=2), da.Grid.make([6, 5, 4], ghost=2, periodic=[1, 0, 0])):
        _raises(lambda: da.fill_values(ctx, A, other, t, phi, rhs, x), "grid differs from the one the CSR structure")

    # host CSR upload: pointers, monotonicity, columns
    B = da.Csr(ctx)
    _raises(lambda: da.upload_host_csr(ctx, B, np.array([0, 1, 3]), np.array([0, 1]), np.array([1.0, 2.0])),
            "row_ptr inconsistent with nnz")
    _raises(lambda: da.upload_host_csr(ctx, B, np.array([0, 2, 1, 3]), np.array([0, 1, 2]),
                                       np.array([1.0, 2.0, 3.0])), "row_ptr not monotone")
    _raises(lambda: da.upload_host_csr(ctx, B, np.array([0, 1, 2]), np.array([0, 5]), np.array([1.0, 2.0])),
            "column index out of range")

    # solver parameters
    da.fill_values(ctx, A, g, t, phi, rhs, x)
    _raises(lambda: da.bicgstab2(ctx, A, rhs, x, precond=4), "precond must be 0 (none), 1 (Jacobi), 2 (multigrid) or 3")
    _raises(lambda: da.refine_mixed(ctx, A, rhs, x, precond=-1), "precond must be 0 (none), 1 (Jacobi) or 2")
    # a vector 8 bytes off a 16-byte boundary (a slice at an odd row): the paired traversal of
    # this even-k grid moves 16-byte vectors, so it is refused before any kernel reads it
    x_odd = torch.zeros(121, dtype=torch.float64, device="cuda")[1:]
    _raises(lambda: da.bicgstab2(ctx, A, rhs, x_odd), "16-byte boundary")
    _raises(lambda: da.refine_mixed(ctx, A, rhs, x_odd), "16-byte boundary")
    _raises(lambda: da.spmv(ctx, A, x, x_odd), "16-byte boundary")
    _raises(lambda: da.residual(ctx, A, rhs, x, x_odd), "16-byte boundary")

    # batch configuration and systems
    Bt = da.Batch(ctx, [A])
    _raises(lambda: Bt.set_config(tile_k=7), "tile_k")
    _raises(lambda: Bt.set_config(grid=12), "grid")
    Bt.close()
    C = da.Csr(ctx)                            # a 2-D system has no diagonal view
    da.build_structure(ctx, C, da.Grid.make([6, 5], ghost=2), da.Terms.make(dt=1.0, diff=1.0, bc_type=[1] * 4))
    _raises(lambda: da.Batch(ctx, [C]), "diagonal view")
    C.close()

    # network shapes
    _raises(lambda: WarmStartNet(ctx, [6, 8, 8], in_channels=3), "dims must be multiples of 4")
    _raises(lambda: WarmStartNet(ctx, [8, 8, 8], in_channels=2), "in_channels")

    # the context still works: a Dirichlet box with a source, solved to 1e-12
    src = torch.ones(g.num_padded, dtype=torch.float64, device="cuda")
    t2 = da.Terms.make(dt=0.5, diff=1.0, source=src, bc_type=[1] * 6, bc_value=[0.0] * 6)
    D = da.Csr(ctx)
    da.build_structure(ctx, D, g, t2)
    da.fill_values(ctx, D, g, t2, phi, rhs, x)
    r = da.bicgstab2(ctx, D, rhs, x, tol=1e-12, tol_kind=1, max_it=200)
    assert r["converged"] == 1
    y = torch.empty_like(x)
    da.spmv(ctx, D, x, y)
    assert float(torch.linalg.norm(rhs - y) / torch.linalg.norm(rhs)) <= 1e-12
    for M in (A, B, D):
        M.close()
