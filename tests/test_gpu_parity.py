"""GPU parity: every HIP kernel, called through the C-ABI, against the CPU oracle.

Stencils, ghost fill, TVD and assembly are bit-exact (the kernels keep the reference's
operation order); the CSR structure is bit-exact; SpMV is bit-exact (same per-row summation
order); BiCGStab2 solutions agree within 1e-12 relative (dot products are reduced in a
different order than the oracle's blocked sum).
"""
import ctypes

import numpy as np
import pytest

import dare_amd as da
import oracle as orc
from conftest import bits_equal

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t, ctx=None):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def grids(g_orc):
    """dare_amd.Grid with the same fields as an oracle grid."""
    g = da.Grid()
    ctypes.memmove(ctypes.byref(g), ctypes.byref(g_orc), ctypes.sizeof(g))
    return g


CASES = [
    dict(n=[12, 13, 14], ncomp=1, periodic=None),
    dict(n=[9, 10, 11], ncomp=3, periodic=[1, 0, 1]),
    dict(n=[17, 15], ncomp=2, periodic=[1, 1]),
    dict(n=[23], ncomp=1, periodic=[1]),
    # two k tiles (one partial), j not a multiple of 4: the marching 3-D kernels' column pieces
    # then start and end inside columns and cross column boundaries
    dict(n=[20, 70, 67], ncomp=1, periodic=[0, 1, 0]),
    # the same with an even k extent (the gradient's two-cells-per-lane form, two k tiles, one
    # partial) and an odd i extent: pieces of odd and even length, two planes per step plus
    # the remainder plane
    dict(n=[21, 70, 130], ncomp=1, periodic=[0, 1, 0]),
]


def make_field(rng, g, flat_frac=0.3):
    npad = orc.lib().orc_num_cells_padded(ctypes.byref(g)) * g.ncomp
    phi = rng.normal(size=npad)
    mask = rng.uniform(size=npad) < flat_frac     # flat patches: 0/0 ratios, inf ratios
    phi[mask] = 1.0
    return phi


def face_vel(rng, g):
    out = []
    for d in range(g.dim):
        u = rng.uniform(-1, 1, size=orc.lib().orc_num_faces(ctypes.byref(g), d))
        u[rng.uniform(size=u.size) < 0.05] = 0.0
        out.append(u)
    return out


@pytest.mark.parametrize("case", range(len(CASES)))
def test_ghost_fill_matches_oracle(ctx, rng, case):
    c = CASES[case]
    go = orc.grid(c["n"], ghost=2, ncomp=c["ncomp"], periodic=c["periodic"])
    dim = go.dim
    modes, values = [], []
    for d in range(dim):
        for side in range(2):
            if c["periodic"] and c["periodic"][d]:
                modes.append(1)
            else:
                modes.append([2, 3, 4, 0][(d + side) % 4])
            values.append(float(rng.normal()))
    phi = make_field(rng, go, 0.0)
    exp = orc.ghost_fill(go, phi, modes, values + [0.0] * (6 - len(values)))
    t = dev(phi)
    da.ghost_fill(ctx, grids(go), t, modes, values + [0.0] * (6 - len(values)))
    assert bits_equal(host(t), exp)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_grad_and_div_faces_match_oracle(ctx, rng, case):
    import torch
    c = CASES[case]
    go = orc.grid(c["n"], size=[0.7, 1.3, 2.1][:len(c["n"])], ghost=2, ncomp=c["ncomp"])
    phi = make_field(rng, go)
    exp = orc.grad_faces(go, phi)
    g = grids(go)
    fd = [torch.zeros(e.size, dtype=torch.float64, device="cuda") for e in exp]
    da.grad_faces(ctx, g, dev(phi), fd)
    for d in range(go.dim):
        assert bits_equal(host(fd[d]), exp[d]), d
    faces = [rng.normal(size=e.size) for e in exp]
    exp_div = orc.div_faces(go, faces)
    out = torch.zeros(exp_div.size, dtype=torch.float64, device="cuda")
    da.div_faces(ctx, g, [dev(f) for f in faces], out)
    assert bits_equal(host(out), exp_div)


@pytest.mark.parametrize("lim", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_tvd_flux_and_step_match_oracle(ctx, rng, lim, case):
    import torch
    c = CASES[case]
    go = orc.grid(c["n"], size=[1.0, 0.8, 1.1][:len(c["n"])], ghost=2, ncomp=c["ncomp"])
    phi = make_field(rng, go)
    vel = face_vel(rng, go)
    exp = orc.tvd_face_flux(go, lim, phi, vel)
    g = grids(go)
    fd = [torch.zeros(e.size, dtype=torch.float64, device="cuda") for e in exp]
    vd = [dev(v) for v in vel]
    pd = dev(phi)
    da.tvd_face_flux(ctx, g, lim, pd, vd, fd)
    for d in range(go.dim):
        assert bits_equal(host(fd[d]), exp[d]), (lim, d)
    dt = 1e-3
    exp_step = orc.tvd_explicit_step(go, lim, phi, vel, dt)
    out = pd.clone()
    da.tvd_explicit_step(ctx, g, lim, pd, vd, dt, out)
    assert bits_equal(host(out), exp_step)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_interpolate_to_faces_and_tvd_with_cell_velocity(ctx, rng, case):
    """Cell-centred velocity fields -> face arrays (the TVD constructor's cell-centred source,
    TVD_Cartesian.inl:27-38) bit-exact against the oracle, then the TVD step driven by them;
    a constant velocity vector (TVD_Cartesian.inl:41-53) as filled face arrays."""
    import torch
    c = CASES[case]
    go = orc.grid(c["n"], size=[1.0, 0.8, 1.1][:len(c["n"])], ghost=2, ncomp=c["ncomp"])
    npad = orc.lib().orc_num_cells_padded(ctypes.byref(go))
    ucell = [rng.uniform(-1, 1, size=npad) for _ in range(go.dim)]
    exp = orc.interpolate_to_faces(go, ucell)
    g = grids(go)
    fd = [torch.zeros(e.size, dtype=torch.float64, device="cuda") for e in exp]
    da.interpolate_to_faces(ctx, g, [dev(u) for u in ucell], fd)
    for d in range(go.dim):
        assert bits_equal(host(fd[d]), exp[d]), d
    phi = make_field(rng, go)
    pd = dev(phi)
    out = pd.clone()
    da.tvd_explicit_step(ctx, g, da.VANLEER, pd, fd, 1e-3, out)
    assert bits_equal(host(out), orc.tvd_explicit_step(go, da.VANLEER, phi, exp, 1e-3))
    const = [0.7, -0.4, 0.25][:go.dim]
    cf = [torch.full((e.size,), const[d], dtype=torch.float64, device="cuda") for d, e in enumerate(exp)]
    da.tvd_explicit_step(ctx, g, da.MINMOD, pd, cf, 1e-3, out)
    ce = [np.full(e.size, const[d]) for d, e in enumerate(exp)]
    assert bits_equal(host(out), orc.tvd_explicit_step(go, da.MINMOD, phi, ce, 1e-3))


ASSEMBLY = ["diff_dirichlet", "var_conv_src", "periodic_conv", "react2", "neumann_2d", "steady_1d", "fixed_x",
            "fixed_react2", "cn_var_conv_src", "ab_periodic_conv", "am_vardiff", "ef_fixed_x"]
# the divergence time schemes (Divergence<Grid, TimeScheme>): a base case under another scheme
SCHEME_CASES = {"cn_var_conv_src": ("var_conv_src", da.CRANK_NICHOLSON),
                "ab_periodic_conv": ("periodic_conv", da.ADAMS_BASHFORT),
                "am_vardiff": ("vardiff_sym", da.ADAMS_MOULTON),
                "ef_fixed_x": ("fixed_x", da.EULER_FORWARD)}


def assembly_case(rng, name):
    if name in SCHEME_CASES:
        base, scheme = SCHEME_CASES[name]
        go, kw, arrays, phi = assembly_case(rng, base)
        kw["div_scheme"] = scheme
        return go, kw, arrays, phi
    if name == "diff_dirichlet":
        go = orc.grid([10, 11, 12], size=[1, 1, 1], ghost=2)
        kw = dict(dt=1e-2, diff=1e-3, bc_type=[1] * 6, bc_value=[0.0] * 6)
    elif name == "var_conv_src":
        go = orc.grid([9, 8, 10], size=[1, 0.9, 1.2], ghost=2)
        kw = dict(dt=5e-3, diff_cell=True, limiter=4, source=True, bc_type=[1, 2, 1, 2, 1, 1],
                  bc_value=[0.3, 0, -1.0, 0, 0.5, 0.25])
    elif name == "periodic_conv":
        go = orc.grid([8, 9, 7], ghost=2, periodic=[1, 1, 1])
        kw = dict(dt=5e-3, diff=2e-3, limiter=3)
    elif name == "periodic_even":     # even k extent: the paired-cell traversal, k wrap
        go = orc.grid([8, 9, 10], ghost=2, periodic=[1, 1, 1])
        kw = dict(dt=5e-3, diff=2e-3, limiter=4)
    elif name == "periodic_sym":      # constant diffusion (exactly symmetric), periodic, k extent 4m:
        # the mixed solve's FP32 quad traversal reads E, N, T from the neighbours' W, S, B slots,
        # across the wrap of every axis
        go = orc.grid([6, 7, 12], ghost=2, periodic=[1, 1, 1])
        kw = dict(dt=5e-3, diff=2e-3)
    elif name == "periodic_single":   # one cell along a periodic x: W and E couple the cell to
        # itself (merged into the CSR diagonal); the FP32 view is then checked for symmetry on
        # the device instead of being taken as symmetric by construction
        go = orc.grid([1, 6, 12], ghost=2, periodic=[1, 0, 0])
        kw = dict(dt=5e-3, diff=2e-3, bc_type=[0, 0, 2, 2, 1, 1], bc_value=[0, 0, 0, 0, 0.5, -0.25])
    elif name == "wide_dirichlet":    # two 128-cell k tiles, the second partial
        go = orc.grid([6, 5, 140], size=[0.5, 0.4, 1.0], ghost=2)
        kw = dict(dt=1e-2, diff=1e-3, bc_type=[1, 2, 2, 1, 1, 1], bc_value=[0.5, 0, 0, 1.0, 0, -0.5])
    elif name == "vardiff_sym":       # per-cell diffusivity (harmonic face means, taken in face
        # order) + source, no convection, no periodic axis: exactly symmetric by construction,
        # applied with the four-slot pair traversal
        go = orc.grid([9, 10, 12], size=[0.9, 1, 1.2], ghost=2)
        kw = dict(dt=5e-3, diff_cell=True, source=True, bc_type=[1, 2, 1, 1, 2, 1],
                  bc_value=[0.25, 0, -0.5, 0.75, 0, 1.0])
    elif name == "periodic_vardiff":  # per-cell diffusivity across periodic wraps: the wrapped
        # faces read the (random) ghost diffusivities, so the device check finds the view
        # asymmetric and the seven-slot traversal runs
        go = orc.grid([6, 7, 10], ghost=2, periodic=[1, 1, 1])
        kw = dict(dt=5e-3, diff_cell=True)
    elif name == "fixed_diff":        # fixed x faces (identity rows) without convection: a
        # candidate for the symmetric traversal that the device check must reject
        go = orc.grid([10, 9, 8], size=[10 / 9, 1, 1], ghost=2)
        kw = dict(dt=5e-3, diff=2e-3, bc_type=[3, 3, 1, 2, 1, 1], bc_value=[0.5, -0.25, 0.1, 0, 0.2, 0.3])
    elif name == "react2":
        go = orc.grid([7, 8, 9], ghost=2, ncomp=2)
        kw = dict(dt=0.1, diff_comp=[1e-3, 5e-3], react=True, bc_type=[1] * 6,
                  bc_value=[0.0] * 6)
    elif name == "fixed_x":           # fixed boundary cells (identity rows) on W/E, one cell longer in x
        go = orc.grid([10, 9, 8], size=[10 / 9, 1, 1], ghost=2)
        kw = dict(dt=5e-3, diff_cell=True, limiter=4, source=True, bc_type=[3, 3, 1, 2, 1, 1],
                  bc_value=[0.5, -0.25, 0.1, 0, 0.2, 0.3])
    elif name == "fixed_react2":      # fixed faces drop the cross-component entry too
        go = orc.grid([7, 8, 9], ghost=2, ncomp=2)
        kw = dict(dt=0.1, diff_comp=[1e-3, 5e-3], react=True, bc_type=[1, 1, 3, 1, 1, 3],
                  bc_value=[0.0, 0.0, 0.75, 0.0, 0.0, -0.5])
    elif name == "neumann_2d":
        go = orc.grid([13, 11], ghost=2, ncomp=3)
        kw = dict(dt=0.2, diff=0.7, limiter=2, bc_type=[2, 1, 1, 2], bc_value=[0, 1.0, 0.5, 0])
    else:
        go = orc.grid([31], ghost=2)
        kw = dict(dt=0.0, diff=1.0, bc_type=[1, 1], bc_value=[0.0, 1.0])
    npad = orc.lib().orc_num_cells_padded(ctypes.byref(go))
    arrays = {}
    if kw.pop("diff_cell", False):
        arrays["diff_cell"] = np.exp(rng.normal(np.log(1e-3), 1.0, size=npad * go.ncomp))
    if kw.pop("source", False):
        arrays["source"] = rng.uniform(size=npad * go.ncomp)
    if kw.pop("react", False):
        ncell = orc.lib().orc_num_cells_internal(ctypes.byref(go))
        k1, k2 = rng.uniform(0.1, 10, ncell), rng.uniform(0.1, 10, ncell)
        arrays["react"] = np.stack([-k1, k2, k1, -k2], 1).ravel().copy()
    if kw.get("limiter", -1) >= 0:
        arrays["vel"] = face_vel(rng, go)
    phi = make_field(rng, go, 0.2)
    return go, kw, arrays, phi


@pytest.mark.parametrize("name", ASSEMBLY)
def test_assembly_matches_oracle(ctx, rng, name):
    import torch
    go, kw, arrays, phi = assembly_case(rng, name)
    to = orc.terms(**kw, **arrays)
    rp, col, val, rhs, x0 = orc.assemble(go, to, phi)
    g = grids(go)
    darr = {k: ([dev(v) for v in a] if isinstance(a, list) else dev(a)) for k, a in arrays.items()}
    td = da.Terms.make(**kw, **darr)
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, td)
    drhs = torch.zeros(len(rhs), dtype=torch.float64, device="cuda")
    dx0 = torch.zeros(len(rhs), dtype=torch.float64, device="cuda")
    da.fill_values(ctx, A, g, td, dev(phi), drhs, dx0)
    ctx.synchronize()
    grp, gcol, gval = A.download()
    assert np.array_equal(grp, rp) and np.array_equal(gcol, col)
    assert bits_equal(gval, val)
    assert bits_equal(host(drhs), rhs) and bits_equal(host(dx0), x0)
    # values are refilled in place on the same structure (BuildReplace semantics)
    phi2 = make_field(rng, go, 0.2)
    rp2, col2, val2, rhs2, _ = orc.assemble(go, to, phi2)
    da.fill_values(ctx, A, g, td, dev(phi2), drhs, dx0)
    ctx.synchronize()
    assert bits_equal(A.download()[2], val2) and bits_equal(host(drhs), rhs2)
    # SpMV: identical per-row summation order -> bit-exact
    x = rng.normal(size=len(rhs))
    y = torch.zeros(len(rhs), dtype=torch.float64, device="cuda")
    da.spmv(ctx, A, dev(x), y)
    assert bits_equal(host(y), orc.spmv(rp2, col2, val2, x))
    A.close()


@pytest.mark.parametrize("name", ["diff_dirichlet", "var_conv_src", "periodic_conv", "react2"])
def test_bicgstab2_matches_oracle(ctx, rng, name):
    import torch
    go, kw, arrays, phi = assembly_case(rng, name)
    to = orc.terms(**kw, **arrays)
    rp, col, val, rhs, x0 = orc.assemble(go, to, phi)
    tol = 1e-12 * np.linalg.norm(rhs)
    xo, ok, it, res, hist = orc.bicgstab2(rp, col, val, rhs, x0, tol, 200, 200)
    assert ok
    A = da.Csr(ctx)
    da.upload_host_csr(ctx, A, rp, col, val)
    xd = dev(x0)
    r = da.bicgstab2(ctx, A, dev(rhs), xd, tol=tol, tol_kind=0, max_it=200, history=200)
    assert r["converged"] == 1 and r["status"] == 0
    assert abs(r["iterations"] - it) <= 1
    xg = host(xd)
    assert np.max(np.abs(xg - xo)) <= 1e-12 * np.max(np.abs(xo)) + tol
    # residual history of the first passes: same method, dots reduced in another order
    k = min(len(hist), len(r["history"]), 3)
    for a, b in zip(r["history"][:k], hist[:k]):
        assert abs(a - b) <= 1e-8 * abs(b) + 1e-14 * float(np.max(hist))
    A.close()


def test_bicgstab2_pseudo2d_laplace_closed_form(ctx):
    from test_oracle import csr_from_dense, laplace_pseudo2d
    Ad, b, x0 = laplace_pseudo2d()
    rp, col, val = csr_from_dense(Ad)
    A = da.Csr(ctx)
    da.upload_host_csr(ctx, A, rp, col, val)
    xd = dev(x0)
    r = da.bicgstab2(ctx, A, dev(b), xd, tol=1e-12, max_it=5000)
    assert r["converged"] == 1
    nodes = np.repeat(np.arange(23), 4)
    assert np.max(np.abs(host(xd) - (nodes + 0.5) / 23)) < 1e-8


def test_bicgstab2_breakdown_and_limits(ctx):
    import torch
    from test_oracle import csr_from_dense, laplace_pseudo2d
    Ad, b, x0 = laplace_pseudo2d()
    rp, col, val = csr_from_dense(Ad)
    A = da.Csr(ctx)
    da.upload_host_csr(ctx, A, rp, col, val)
    xd = dev(x0)
    r = da.bicgstab2(ctx, A, dev(b), xd, tol=1e3)
    assert r["converged"] == 1 and r["iterations"] == 0
    z = torch.zeros(len(b), dtype=torch.float64, device="cuda")
    r = da.bicgstab2(ctx, A, z, z.clone(), tol=1e-12)
    assert r["converged"] == 1 and r["iterations"] == 0
    xd = dev(x0)
    r = da.bicgstab2(ctx, A, dev(b), xd, tol=1e-30, max_it=3)
    assert r["converged"] == 0 and r["iterations"] == 4
    # zero operator: gamma == 0 breakdown is reported, not raised
    Z = da.Csr(ctx)
    da.upload_host_csr(ctx, Z, rp, col, np.zeros_like(val))
    r = da.bicgstab2(ctx, Z, dev(b), dev(x0), tol=1e-12)
    assert r["converged"] == 0 and r["status"] == da.DARE_ERR_BREAKDOWN


@pytest.mark.parametrize("name", ["diff_dirichlet", "var_conv_src"])
def test_mixed_refinement_reaches_fp64_residual(ctx, rng, name):
    go, kw, arrays, phi = assembly_case(rng, name)
    to = orc.terms(**kw, **arrays)
    rp, col, val, rhs, x0 = orc.assemble(go, to, phi)
    A = da.Csr(ctx)
    da.upload_host_csr(ctx, A, rp, col, val)
    xd = dev(x0)
    r = da.refine_mixed(ctx, A, dev(rhs), xd, tol=1e-10, tol_kind=1, inner_tol=1e-4)
    assert r["converged"] == 1
    x = host(xd)
    true_rel = np.linalg.norm(rhs - orc.spmv(rp, col, val, x)) / np.linalg.norm(rhs)
    assert true_rel <= 1e-10
    xo, ok, *_ = orc.bicgstab2(rp, col, val, rhs, x0, 1e-13 * np.linalg.norm(rhs), 300)
    assert np.max(np.abs(x - xo)) <= 1e-8 * np.max(np.abs(xo))


DIA_CASES = ["diff_dirichlet", "var_conv_src", "periodic_conv", "periodic_even", "periodic_sym", "periodic_single",
             "wide_dirichlet", "react2", "fixed_x", "fixed_react2", "vardiff_sym", "periodic_vardiff", "fixed_diff",
             "cn_var_conv_src", "am_vardiff"]


def device_assembly(ctx, rng, name):
    """Assemble on the device (structure + values): the matrix then carries the diagonal view
    the solver traverses."""
    import torch
    go, kw, arrays, phi = assembly_case(rng, name)
    to = orc.terms(**kw, **arrays)
    ref = orc.assemble(go, to, phi)
    g = grids(go)
    darr = {k: ([dev(v) for v in a] if isinstance(a, list) else dev(a)) for k, a in arrays.items()}
    td = da.Terms.make(**kw, **darr)
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, td)
    rhs = torch.zeros(len(ref[3]), dtype=torch.float64, device="cuda")
    x0 = torch.zeros_like(rhs)
    da.fill_values(ctx, A, g, td, dev(phi), rhs, x0)
    ctx.synchronize()
    return go, A, rhs, x0, ref, (td, darr)


@pytest.mark.parametrize("name", DIA_CASES)
def test_diagonal_view_spmv_matches_oracle(ctx, rng, name):
    """The solver's operator (tiled traversal of the diagonal view) against the oracle's CSR
    SpMV: same per-row order, so bit-exact except on periodic-wrapped rows."""
    import torch
    go, A, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    x = rng.normal(size=len(orhs))
    w = rng.normal(size=len(orhs))
    y = torch.zeros(len(orhs), dtype=torch.float64, device="cuda")
    d = da.spmv_dot(ctx, A, dev(x), y, dev(w))
    exp = orc.spmv(rp, col, val, x)
    got = host(y)
    if name.startswith("periodic"):
        assert np.max(np.abs(got - exp)) <= 1e-14 * np.max(np.abs(exp))
    else:
        assert bits_equal(got, exp)
    assert abs(d - float(exp @ w)) <= 1e-12 * float(np.abs(exp) @ np.abs(w))
    A.close()


CONVECTION_CASES = {"var_conv_src", "periodic_conv", "periodic_even", "fixed_x", "cn_var_conv_src"}


def csr_bitwise_symmetric(rp, col, val):
    """A == A^T entry for entry and bit for bit (structure included)."""
    import scipy.sparse as sp
    n = len(rp) - 1
    M = sp.csr_matrix((val, col, rp), shape=(n, n))
    T = M.T.tocsr()
    M.sort_indices()
    T.sort_indices()
    return (np.array_equal(M.indptr, T.indptr) and np.array_equal(M.indices, T.indices)
            and np.array_equal(M.data.view(np.uint64), T.data.view(np.uint64)))


@pytest.mark.parametrize("name", DIA_CASES)
def test_symmetric_traversal_only_on_symmetric_operators(ctx, rng, name):
    """The four-slot symmetric traversal runs exactly when the operator is symmetric: the
    solver's flag (by construction of the fill, or the device's slot-by-slot check) against
    the oracle's CSR, A == A^T bit for bit, for the one-component operators without convection
    whose k extent is even (the paired traversal); never for the others."""
    go, A, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    eligible = go.ncomp == 1 and go.n[2] % 2 == 0 and name not in CONVECTION_CASES
    expected = eligible and csr_bitwise_symmetric(rp, col, val)
    assert da.csr_symmetric(ctx, A) == expected
    if name in ("diff_dirichlet", "vardiff_sym", "periodic_sym", "wide_dirichlet"):
        assert expected       # the cases that exercise the symmetric traversal
    if name in ("periodic_vardiff", "fixed_diff"):
        assert eligible and not expected
    A.close()


@pytest.mark.parametrize("name", ["diff_dirichlet", "vardiff_sym", "periodic_sym", "wide_dirichlet"])
def test_symmetric_traversal_is_bit_identical(ctx, rng, name):
    """Four-slot (symmetric) against seven-slot traversal on the same matrix: SpMV + dot, the
    FP64 solve (FP64 pair traversal) and the mixed solve (FP32 quad traversal, FP64
    residuals) give the same bits and the same pass counts."""
    import torch
    go, A, rhs, x0, ref, keep = device_assembly(ctx, rng, name)
    n = len(ref[3])
    x = dev(rng.normal(size=n))
    w = dev(rng.normal(size=n))
    out = {}
    for on in (True, False):
        da.set_symmetric_traversal(A, on)
        assert da.csr_symmetric(ctx, A) == on
        y = torch.zeros_like(x)
        d = da.spmv_dot(ctx, A, x, y, w)
        xs = x0.clone()
        r = da.bicgstab2(ctx, A, rhs, xs, tol=1e-10, tol_kind=1, max_it=300)
        xm = x0.clone()
        rm = da.refine_mixed(ctx, A, rhs, xm, tol=1e-10, tol_kind=1, inner_tol=1e-4)
        assert r["converged"] == 1 and rm["converged"] == 1
        out[on] = (host(y), d, host(xs), r["iterations"], host(xm), rm["iterations"], rm["refinements"])
    a, b = out[True], out[False]
    assert bits_equal(a[0], b[0]) and a[1] == b[1]
    assert bits_equal(a[2], b[2]) and a[3] == b[3]
    assert bits_equal(a[4], b[4]) and a[5:] == b[5:]
    A.close()


def _fill_sym_case(ctx, rng_seed, name, stale):
    """A symmetric-by-construction case assembled on the device. stale: the matrix is first
    filled with all slots from a different operator (diffusivity x3), then refilled with the
    case's operator with the symmetric traversal on, which leaves the E, N, T slots out (they
    keep the other operator's values). Otherwise the case is filled with every slot."""
    import torch
    go, kw, arrays, phi = assembly_case(np.random.default_rng(rng_seed), name)
    g = grids(go)
    darr = {k: dev(a) for k, a in arrays.items()}
    td = da.Terms.make(**kw, **darr)
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, td)
    n = orc.lib().orc_num_cells_internal(ctypes.byref(go)) * go.ncomp
    rhs = torch.zeros(n, dtype=torch.float64, device="cuda")
    x0 = torch.zeros_like(rhs)
    da.set_symmetric_traversal(A, False)            # every slot written
    if stale:
        kw2 = dict(kw)
        arr2 = dict(darr)
        if "diff_cell" in arr2:
            arr2["diff_cell"] = arr2["diff_cell"] * 3.0
        else:
            kw2["diff"] = kw["diff"] * 3.0
        da.fill_values(ctx, A, g, da.Terms.make(**kw2, **arr2), dev(phi), rhs, x0)
        da.set_symmetric_traversal(A, True)         # the refill leaves E, N, T out
        da.fill_values(ctx, A, g, td, dev(phi), rhs, x0)
    else:
        da.fill_values(ctx, A, g, td, dev(phi), rhs, x0)
        da.set_symmetric_traversal(A, True)
    ctx.synchronize()
    return A, rhs, x0, (g, td, darr)


LAZY_CONSUMERS = ["csr", "spmv_seven_slot", "jacobi", "jacobi_left", "mg", "mixed_jacobi", "mixed_mg", "batch"]


@pytest.mark.parametrize("consumer", LAZY_CONSUMERS)
@pytest.mark.parametrize("name", ["diff_dirichlet", "vardiff_sym", "periodic_sym"])
def test_consumers_of_a_view_filled_without_east_north_top(ctx, name, consumer):
    """An operator that is symmetric by construction is filled without its E, N, T slots (the
    symmetric traversals read them as the neighbours' W, S, B slots). Every consumer that reads
    the seven slots -- CSR values, the seven-slot traversal, Jacobi scaling on either side,
    multigrid, the FP32 copy's scaled view, batches -- must see them written: the slots first
    hold another operator's values, and each result is compared bit for bit with the same call
    on the same operator filled with every slot."""
    import torch
    out = []
    for stale in (True, False):
        A, rhs, x0, keep = _fill_sym_case(ctx, 7, name, stale)
        assert da.csr_symmetric(ctx, A)
        n = rhs.numel()
        if consumer == "csr":
            got = A.download()[2]
        elif consumer == "spmv_seven_slot":
            da.set_symmetric_traversal(A, False)
            x = dev(np.random.default_rng(3).normal(size=n))
            y = torch.zeros_like(x)
            da.spmv_dot(ctx, A, x, y, x)
            got = host(y)
        elif consumer in ("jacobi", "jacobi_left", "mg"):
            pre = {"jacobi": 1, "jacobi_left": 3, "mg": 2}[consumer]
            x = x0.clone()
            r = da.bicgstab2(ctx, A, rhs, x, tol=1e-10, tol_kind=1, max_it=300, precond=pre)
            assert r["converged"] == 1
            got = (host(x), r["iterations"])
        elif consumer in ("mixed_jacobi", "mixed_mg"):
            x = x0.clone()
            r = da.refine_mixed(ctx, A, rhs, x, tol=1e-10, tol_kind=1, inner_tol=1e-4,
                                precond=1 if consumer == "mixed_jacobi" else 2)
            assert r["converged"] == 1
            got = (host(x), r["iterations"], r["refinements"])
        else:
            B = da.Batch(ctx, [A])
            for src, addr in ((host(rhs), B.rhs_ptr(0)), (host(x0), B.x_ptr(0))):
                da.check(da.lib().dare_memcpy_h2d(ctx.h, addr, src.ctypes.data, 8 * n), "h2d")
            r = B.solve(tol=1e-10, tol_kind=1, max_it=300)[0]
            assert r["converged"] == 1
            x = np.empty(n)
            da.check(da.lib().dare_memcpy_d2h(ctx.h, x.ctypes.data, B.x_ptr(0), 8 * n), "d2h")
            ctx.synchronize()
            got = (x, r["iterations"])
            B.close()
        out.append(got)
        A.close()
    a, b = out
    if isinstance(a, tuple):
        assert bits_equal(a[0], b[0]) and a[1:] == b[1:]
    else:
        assert bits_equal(a, b)


@pytest.mark.parametrize("name", DIA_CASES)
def test_solver_on_device_assembly_matches_oracle(ctx, rng, name):
    go, A, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    tol = 1e-12 * np.linalg.norm(orhs)
    xo, ok, it, res, hist = orc.bicgstab2(rp, col, val, orhs, ox0, tol, 300)
    assert ok
    r = da.bicgstab2(ctx, A, rhs, x0, tol=tol, tol_kind=0, max_it=300)
    assert r["converged"] == 1 and abs(r["iterations"] - it) <= 1
    x = host(x0)
    assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo)) + tol
    A.close()


@pytest.mark.parametrize("name", DIA_CASES)
def test_mixed_on_device_assembly(ctx, rng, name):
    go, A, rhs, x0, (rp, col, val, orhs, ox0), keep = device_assembly(ctx, rng, name)
    r = da.refine_mixed(ctx, A, rhs, x0, tol=1e-10, tol_kind=1, inner_tol=1e-4)
    assert r["converged"] == 1
    x = host(x0)
    assert np.linalg.norm(orhs - orc.spmv(rp, col, val, x)) <= 1e-10 * np.linalg.norm(orhs)
    A.close()


def test_zero_guess_inner_solve_that_runs_no_pass_returns_zero(ctx, rng):
    """The inner FP32 solves start from a zero guess whose initialisation writes no vector (the
    first pass writes x and r). An inner solve that stops before its first pass (here: an inner
    tolerance above 1, so the zero guess already meets it) must still hand back x = 0: after a
    solve that left other values in the work vectors, refinements with such inner solves leave
    the iterate bit for bit unchanged."""
    go, A, rhs, x0, ref, keep = device_assembly(ctx, rng, "diff_dirichlet")
    warm = x0.clone()
    assert da.refine_mixed(ctx, A, rhs, warm, tol=1e-10, tol_kind=1, inner_tol=1e-4)["converged"] == 1
    x = dev(rng.normal(size=x0.numel()))
    before = host(x).copy()
    r = da.refine_mixed(ctx, A, rhs, x, tol=1e-10, tol_kind=1, inner_tol=2.0, max_refine=3)
    assert r["converged"] == 0 and r["refinements"] == 3 and r["iterations"] == 0
    assert bits_equal(host(x), before)
    A.close()


def test_full_size_structure_and_solve_properties(ctx):
    """256^3 (the benchmark size), properties that do not need the oracle: nnz = 7n^3 - 6n^2,
    row lengths 4..7, sorted columns, SpMV of ones = row sums, and the FP64 solve's true
    relative residual."""
    import torch
    from dare_amd.workloads import C2Rotation
    n = 256
    W = C2Rotation(n)
    A = da.Csr(ctx)
    da.build_structure(ctx, A, W.grid, W.terms)
    da.fill_values(ctx, A, W.grid, W.terms, W.phi, W.rhs, W.x)
    ctx.synchronize()
    nr, nnz, rp_ptr, col_ptr, val_ptr = A.info()
    assert nr == n ** 3 and nnz == 7 * n ** 3 - 6 * n ** 2
    rp, col, val = A.download()
    lens = np.diff(rp)
    assert lens.min() == 4 and lens.max() == 7 and rp[-1] == nnz
    first = rp[:-1]
    d = np.diff(col)
    start_mask = np.zeros(nnz, bool)
    start_mask[first] = True
    assert np.all(d[~start_mask[1:]] > 0)        # ascending within every row
    ones = torch.ones(nr, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(ones)
    da.spmv(ctx, A, ones, y)
    rowsum = np.add.reduceat(val, first)
    assert np.array_equal(host(y), rowsum) or np.allclose(host(y), rowsum, rtol=1e-15, atol=0)
    r = da.bicgstab2(ctx, A, W.rhs, W.x, tol=1e-8, tol_kind=1, max_it=100)
    assert r["converged"] == 1
    da.spmv(ctx, A, W.x, y)
    rel = float(torch.linalg.norm(W.rhs - y) / torch.linalg.norm(W.rhs))
    assert rel <= 1e-8
    A.close()
