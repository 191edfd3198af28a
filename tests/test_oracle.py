"""Pins the CPU oracle (oracle/oracle.cpp) before the GPU results are checked against it.

Each check is one the reference itself holds or one the mathematics fixes:
  * the reference's own headers compiled into oracle/_ref (limiters, Diffusion_N0D1, the
    utils::Vector operator order of the TVD face expressions),
  * the known-answer tests of test_Cartesian_Operators.cpp, test_DDT.cpp, test_HaloBuffer.cpp,
  * closed-form solves of test_TrilinosSolver.cpp / test_integration_Cartesian_Trilinos.cpp,
  * dense LU on small grids, Diffusion_N0D1 as a time-dependent verification.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle as orc
from conftest import bits_equal

LIMS = [0, 1, 2, 3]   # CDS, UPWIND, VANALBADA, MINMOD (the reference's four)


def padded_index(g, ind):
    o = 0
    for d in range(g.dim):
        s = 1
        for e in range(d + 1, g.dim):
            s *= g.n[e] + 2 * g.ghost
        o += s * ind[d]
    return o


def face_flat(g, axis, q):
    nf = [g.n[e] + (1 if e == axis else 0) for e in range(g.dim)]
    o = 0
    for d in range(g.dim):
        s = 1
        for e in range(d + 1, g.dim):
            s *= nf[e]
        o += s * q[d]
    return o


def field_from(g, fn):
    """Padded field with value fn(padded index tuple, component)."""
    shape = [g.n[d] + 2 * g.ghost for d in range(g.dim)]
    out = np.zeros(int(np.prod(shape)) * g.ncomp)
    for flat, idx in enumerate(np.ndindex(*shape)):
        for c in range(g.ncomp):
            out[flat * g.ncomp + c] = fn(idx, c)
    return out


def geometry(g):
    dx = [g.size[d] / g.n[d] for d in range(g.dim)]
    V = 1.0
    for d in range(g.dim):
        V *= dx[d]
    A = [V / dx[d] for d in range(g.dim)]
    dnr = [1.0 / dx[d] for d in range(g.dim)]
    return dx, V, A, dnr


# ------------------------------------------------------------------ reference headers (_ref)

SPECIAL_R = [0.0, -0.0, 1.0, -1.0, 0.5, 2.0, 1e-300, -1e-300, 1e300, -1e300, math.inf, -math.inf,
             math.nan, 3.0, -7.25]


@pytest.mark.parametrize("lim", LIMS)
def test_limiters_match_reference_headers(lim, rng):
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rs = SPECIAL_R + list(rng.normal(0, 3, 200)) + list(10.0 ** rng.uniform(-8, 8, 50))
    for r in rs:
        a = orc.lib().orc_limiter(lim, r)
        b = R.ref_limiter(lim, r)
        assert bits_equal(np.array([a]), np.array([b])), (lim, r, a, b)


def test_diffusion_n0d1_matches_reference(rng):
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    for tau, zeta in zip(rng.uniform(1e-4, 2.0, 100), rng.uniform(0, 1, 100)):
        a = orc.lib().orc_diffusion_n0d1(tau, zeta, 100)
        b = R.ref_diffusion_n0d1(tau, zeta, 100)
        assert a == b


def test_diffusion_n0d1_closed_form_properties():
    f = lambda t, z: orc.lib().orc_diffusion_n0d1(t, z, 100)
    assert abs(f(0.3, 1.0) - 1.0) < 1e-12            # Dirichlet 1 at zeta = 1
    assert abs(f(20.0, 0.3) - 1.0) < 1e-12           # steady state
    h = 1e-6                                          # Neumann at zeta = 0
    assert abs((f(0.2, h) - f(0.2, 0.0)) / h) < 1e-4
    # one-mode limit: 1 - (4/pi) exp(-pi^2 tau / 4) cos(pi zeta / 2) for large tau
    tau = 1.5
    exact1 = 1 - 4 / math.pi * math.exp(-math.pi ** 2 * tau / 4) * math.cos(math.pi * 0.4 / 2)
    assert abs(f(tau, 0.4) - exact1) < 1e-9


def test_tvd_face_expressions_match_reference_vector_ops(rng):
    """1-D grid, constant +/-1 velocity: flux / u and the matrix RHS equal the reference's
    Vector-operator expressions (TVD_Cartesian.inl:152-157, 215-219) bit for bit."""
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    g = orc.grid([6], ghost=2)
    for lim in LIMS:
        for trial in range(20):
            phi = rng.normal(size=10)
            if trial % 4 == 0:
                phi[3:6] = 1.0            # flat region: 0/0 -> NaN -> limiter of NaN
            for u in (1.0, -1.0):
                ux = np.full(7, u)
                fx, = orc.tvd_face_flux(g, lim, phi, [ux])
                for cell in range(6):
                    p = cell + 2
                    if u >= 0:
                        uu, uc, dd = phi[p - 2], phi[p - 1], phi[p]
                    else:
                        uu, uc, dd = phi[p + 1], phi[p], phi[p - 1]
                    face = fx[cell] * u
                    exp = R.ref_tvd_face_value(lim, uu, uc, dd)
                    assert bits_equal(np.array([face]), np.array([exp])), (lim, cell, u)
                    nb, c, rhs = orc.tvd_matrix_cell(g, lim, phi, [ux], cell)
                    exp_rhs = R.ref_tvd_face_rhs(lim, uu, uc, dd, u)
                    assert bits_equal(np.array([rhs[0]]), np.array([exp_rhs]))


def test_vanleer_closed_form():
    f = lambda r: orc.lib().orc_limiter(4, r)
    assert f(-3.0) == 0.0 and f(0.0) == 0.0 and f(1.0) == 1.0
    assert f(math.nan) == 0.0 and f(math.inf) == 2.0 and f(-math.inf) == 0.0
    for r in np.linspace(0.01, 50, 500):
        psi = f(r)
        assert abs(psi - 2 * r / (1 + r)) < 1e-14 * (1 + psi)     # (r+|r|)/(1+|r|) for r>0
        assert 0 <= psi <= min(2 * r, 2.0) + 1e-15                 # TVD region
        assert abs(psi / r - f(1 / r)) < 1e-13                      # symmetry psi(r)/r = psi(1/r)


# ------------------------------------------------------------------ test_Cartesian_Operators KATs

def kat_grid(dim, ncomp=3):
    return orc.grid([10 + d for d in range(dim)], size=[1.0 + d for d in range(dim)], ghost=2,
                    ncomp=ncomp)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_gradient_face_values_kat(dim):
    """test_Cartesian_Operators.cpp:255-389 (EXPECT_EQ): field = (last index) * n."""
    g = kat_grid(dim)
    phi = field_from(g, lambda idx, c: idx[dim - 1] * c)
    faces = orc.grad_faces(g, phi)
    dx, V, A, dnr = geometry(g)
    p = [2] * dim                      # internal ordinal 0
    q = [0] * dim
    for c in range(3):
        at = lambda ind: phi[padded_index(g, ind) * 3 + c]
        for d in range(dim):
            lo = face_flat(g, d, q)
            qh = list(q)
            qh[d] += 1
            hi = face_flat(g, d, qh)
            pl, ph = list(p), list(p)
            pl[d] -= 1
            ph[d] += 1
            exp_w = (at(p) - at(pl)) * dnr[d]
            exp_e = (at(ph) - at(p)) * dnr[d]
            assert faces[d][lo * 3 + c] == exp_w
            assert faces[d][hi * 3 + c] == exp_e
            if d != dim - 1:
                assert faces[d][lo * 3 + c] == 0.0


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_divergence_face_value_kat(dim):
    """test_Cartesian_Operators.cpp:512-620 (EXPECT_EQ)."""
    g = kat_grid(dim)
    vals = {0: (0.0, 2.0), 1: (3.0, 4.0), 2: (5.0, 6.0)}
    faces = orc.faces_alloc(g)
    q = [0] * dim
    for d in range(dim):
        qh = list(q)
        qh[d] += 1
        for c in range(3):
            faces[d][face_flat(g, d, q) * 3 + c] = vals[d][0] * c
            faces[d][face_flat(g, d, qh) * 3 + c] = vals[d][1] * c
    out = orc.div_faces(g, faces)
    _, _, A, _ = geometry(g)
    lp = padded_index(g, [2] * dim)
    for c in range(3):
        exp = (vals[0][1] - vals[0][0]) * c * A[0]
        if dim > 1:
            exp = exp + (vals[1][1] - vals[1][0]) * c * A[1]
        if dim > 2:
            exp = exp + (vals[2][1] - vals[2][0]) * c * A[2]
        assert out[lp * 3 + c] == exp


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_matrix_block_integration_kat(dim):
    """mb = -1 * div(grad(mb)): W = E = -A dn_r, C = 2 A0 dn0 + ... (EXPECT_EQ,
    test_Cartesian_Operators.cpp:713-760, 939-965)."""
    g = kat_grid(dim)
    t = orc.terms(dt=0.0, diff=1.0)
    phi = np.zeros(orc.lib().orc_num_cells_padded(ctypes.byref(g)) * 3)
    rp, col, val, rhs, x0 = orc.assemble(g, t, phi)
    _, _, A, dnr = geometry(g)
    q = [5] * dim
    node = 0
    for d in range(dim):
        s = 1
        for e in range(d + 1, dim):
            s *= g.n[e]
        node += s * q[d]
    for c in range(3):
        row = node * 3 + c
        cols = col[rp[row]:rp[row + 1]]
        vals = val[rp[row]:rp[row + 1]]
        assert len(cols) == 2 * dim + 1 and np.all(np.diff(cols) > 0)
        center = dict(zip(cols, vals))[row]
        exp_c = 2. * A[0] * dnr[0]
        if dim > 1:
            exp_c = exp_c + 2. * A[1] * dnr[1]
        if dim > 2:
            exp_c = exp_c + 2. * A[2] * dnr[2]
        assert center == exp_c
        # neighbors, in ascending column order: W,S,B,C,T,N,E (last index fastest)
        nb = [v for cc, v in zip(cols, vals) if cc != row]
        exp_nb = [-A[d] * dnr[d] for d in range(dim)] + [-A[d] * dnr[d] for d in reversed(range(dim))]
        assert nb == exp_nb


@pytest.mark.parametrize("lim,quad", [(0, False), (1, False), (3, True), (2, False), (2, True), (3, False)])
def test_tvd_1d_kats(lim, quad):
    """TVD value and matrix KATs (test_Cartesian_Operators.cpp:967-1501)."""
    g = kat_grid(1)
    phi = field_from(g, lambda idx, c: (idx[0] * idx[0] if quad else idx[0]) * c)
    eps = np.finfo(float).eps
    p = 2
    for vel in (1.0, -1.0):
        ux = np.full(11, vel)
        fx, = orc.tvd_face_flux(g, lim, phi, [ux])
        for c in range(3):
            at = lambda i: phi[i * 3 + c]
            pw, pc, pe, pww, pee = at(p - 1), at(p), at(p + 1), at(p - 2), at(p + 2)

            def limf(r):
                if lim == 0:
                    return 1.0
                if lim == 1:
                    return 0.0
                if math.isnan(r):
                    r = 0.0
                if lim == 3:
                    return max(0., min(r, 1.))
                return (r + r * r) / (1. + r * r)

            def ratio(a, b):
                with np.errstate(all="ignore"):
                    return float(np.float64(a) / np.float64(b))

            if vel > 0:
                r_w, r_e = ratio(pw - pww, pc - pw), ratio(pc - pw, pe - pc)
                v_w = pw + 0.5 * limf(r_w) * (pc - pw)
                v_e = pc + 0.5 * limf(r_e) * (pe - pc)
                rhs_w = -0.5 * limf(r_w) * (pc - pw) * vel
                rhs_e = -0.5 * limf(r_e) * (pe - pc) * vel
            else:
                r_w, r_e = ratio(pe - pc, pc - pw), ratio(pee - pe, pe - pc)
                v_w = pc + 0.5 * limf(r_w) * (pw - pc)
                v_e = pe + 0.5 * limf(r_e) * (pc - pe)
                rhs_w = -0.5 * limf(r_w) * (pw - pc) * vel
                rhs_e = -0.5 * limf(r_e) * (pc - pe) * vel
            face_w = fx[0 * 3 + c] * vel
            face_e = fx[1 * 3 + c] * vel
            assert abs(face_w - v_w) <= 1e3 * eps * abs(v_w)
            assert abs(face_e - v_e) <= 1e3 * eps * abs(v_e)
            nb, cc, rhs = orc.tvd_matrix_cell(g, lim, phi, [ux], 0, c)
            if vel > 0:
                assert (nb[0], cc[0], nb[1], cc[1]) == (vel, 0.0, 0.0, vel)
            else:
                assert (nb[0], cc[0], nb[1], cc[1]) == (0.0, vel, vel, 0.0)
            if lim in (1, 2, 3) and quad or lim == 1:
                assert rhs[0] == rhs_w and rhs[1] == rhs_e       # EXPECT_EQ in the reference
            else:
                assert abs(rhs[0] - rhs_w) <= 1e3 * eps * abs(rhs_w)
                assert abs(rhs[1] - rhs_e) <= 1e3 * eps * abs(rhs_e)


def test_ddt_kat():
    """test_DDT.cpp:73-110 EulerBackward s_1: centre V/dt, rhs V/dt * phi_old."""
    g = orc.grid([10, 11, 12], size=[1., 2., 3.], ghost=2, ncomp=3)
    phi_old = np.full(orc.lib().orc_num_cells_padded(ctypes.byref(g)) * 3, 2. * 0.3)
    t = orc.terms(dt=0.1, diff=0.0)
    rp, col, val, rhs, x0 = orc.assemble(g, t, phi_old)
    _, V, _, _ = geometry(g)
    dV_dt = V / 0.1
    eps = np.finfo(float).eps
    for row in range(0, len(rhs), 97):
        cols = col[rp[row]:rp[row + 1]]
        center = val[rp[row]:rp[row + 1]][list(cols).index(row)]
        assert abs(center - dV_dt) <= eps
        assert abs(rhs[row] - dV_dt * 2. * 0.3) <= eps
    assert np.all(x0 == 0.6)


def test_periodic_ghost_fill_halo_buffer_rule():
    """test_HaloBuffer.cpp:33-115 on one rank: ghosts take the wrapped internal values."""
    g = orc.grid([2], ghost=1, ncomp=5, periodic=[1])
    phi = np.full(4 * 5, -1.0)
    for n in range(1, 3):
        for e in range(5):
            phi[n * 5 + e] = n * 5 + e
    out = orc.ghost_fill(g, phi, [1, 1])
    for e in range(5):
        assert out[0 * 5 + e] == (2 * 1 + 0) * 5 + e   # rank 0 low ghost <- last internal
        assert out[3 * 5 + e] == 1 * 5 + e             # high ghost <- first internal


def test_periodic_ghost_fill_3d_faces_only(rng):
    g = orc.grid([4, 5, 6], ghost=2, periodic=[1, 1, 1])
    shape = (8, 9, 10)
    phi = rng.normal(size=int(np.prod(shape)))
    out = orc.ghost_fill(g, phi, [1] * 6).reshape(shape)
    src = phi.reshape(shape)
    for idx in np.ndindex(*shape):
        internal = [0 <= idx[d] - 2 < g.n[d] for d in range(3)]
        nout = 3 - sum(internal)
        if nout == 1:
            d = internal.index(False)
            w = list(idx)
            w[d] = (idx[d] - 2) % g.n[d] + 2
            assert out[idx] == src[tuple(w)]
        else:
            assert out[idx] == src[idx]      # internal cells, edges and corners untouched


# ------------------------------------------------------------------ CSR structure

def test_csr_structure_dirichlet_box():
    n = 9
    g = orc.grid([n, n, n], ghost=1)
    t = orc.terms(dt=1.0, diff=1.0, bc_type=[1] * 6)
    phi = np.zeros((n + 2) ** 3)
    rp, col, val, rhs, x0 = orc.assemble(g, t, phi)
    assert rp[-1] == 7 * n ** 3 - 6 * n ** 2
    for row in range(n ** 3):
        c = col[rp[row]:rp[row + 1]]
        assert np.all(np.diff(c) > 0)
        i, j, k = row // (n * n), (row // n) % n, row % n
        exp = []
        for (di, dj, dk) in [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (0, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 0)]:
            a, b, cc = i + di, j + dj, k + dk
            if 0 <= a < n and 0 <= b < n and 0 <= cc < n:
                exp.append((a * n + b) * n + cc)
        assert list(c) == exp


def test_csr_periodic_wrap_sorted():
    n = 5
    g = orc.grid([n, n, n], ghost=2, periodic=[1, 1, 1])
    t = orc.terms(dt=1.0, diff=0.3)
    rp, col, val, rhs, x0 = orc.assemble(g, t, np.zeros((n + 4) ** 3))
    assert rp[-1] == 7 * n ** 3
    A = np.zeros((n ** 3, n ** 3))
    for r in range(n ** 3):
        c = col[rp[r]:rp[r + 1]]
        assert np.all(np.diff(c) > 0)
        A[r, c] = val[rp[r]:rp[r + 1]]
    assert np.allclose(A, A.T, rtol=0, atol=1e-15)       # symmetric diffusion operator
    assert np.allclose(A.sum(axis=1), 1.0 / n ** 3 / 1.0, rtol=1e-12)   # rows sum to V/dt


# ------------------------------------------------------------------ BiCGStab2

def csr_from_dense(A):
    rp = [0]
    cols, vals = [], []
    for r in range(A.shape[0]):
        nz = np.nonzero(A[r])[0]
        cols += list(nz)
        vals += list(A[r, nz])
        rp.append(len(cols))
    return np.array(rp, np.int32), np.array(cols, np.int32), np.array(vals)


def laplace_pseudo2d(N=4, size=23):
    """Matrix, rhs, x0 of the functor in test_TrilinosSolver.cpp:98-160."""
    n = size * N
    A = np.zeros((n, n))
    b = np.zeros(n)
    x0 = np.full(n, 0.5)
    for node in range(size):
        for c in range(N):
            r = node * N + c
            if node == 0:
                A[r, r] = 3.0
                A[r, r + N] = -1.0
                x0[r] = 0.0
            elif node == size - 1:
                A[r, r - N] = -1.0
                A[r, r] = 3.0
                b[r] = 2.0
                x0[r] = 1.0
            else:
                A[r, r - N] = -1.0
                A[r, r] = 2.0
                A[r, r + N] = -1.0
            if c != 0:
                A[r, r - 1] = -1.0
                A[r, r] += 1
            if c != N - 1:
                A[r, r + 1] = -1.0
                A[r, r] += 1
    return A, b, x0


def test_bicgstab2_pseudo2d_laplace_closed_form():
    A, b, x0 = laplace_pseudo2d()
    rp, col, val = csr_from_dense(A)
    x, ok, it, res, _ = orc.bicgstab2(rp, col, val, b, x0, 1e-12, 5000)
    assert ok
    nodes = np.repeat(np.arange(23), 4)
    assert np.max(np.abs(x - (nodes + 0.5) / 23)) < 1e-8


def test_bicgstab2_1d_linear_profile():
    """test_integration_Cartesian_Trilinos.cpp:82-160 SolveScalar: 20 cells, N = 3."""
    n, N = 20, 3
    A = np.zeros((n * N, n * N))
    b = np.zeros(n * N)
    for i in range(n):
        for c in range(N):
            r = i * N + c
            A[r, r] = 2.0
            if i > 0:
                A[r, r - N] = -1.0
            else:
                A[r, r] += 1.0
                b[r] = 0.0
            if i < n - 1:
                A[r, r + N] = -1.0
            else:
                A[r, r] += 1.0
                b[r] = 2.0
    rp, col, val = csr_from_dense(A)
    x, ok, it, res, _ = orc.bicgstab2(rp, col, val, b, np.zeros(n * N), 1e-14, 5000)
    assert ok
    exp = (np.repeat(np.arange(n), N) + 0.5) / n
    assert np.max(np.abs(x - exp)) < 1e-8


def test_bicgstab2_3d_linear_profile_term_list():
    """3-D linear profile, Dirichlet 0/1 by mirror on W/E, Neumann elsewhere
    (test_integration_Cartesian_Trilinos.cpp 3-D scalar pattern)."""
    g = orc.grid([20, 21, 22], size=[1., 2., 3.], ghost=2)
    t = orc.terms(dt=0.0, diff=1.0, bc_type=[1, 1, 2, 2, 2, 2], bc_value=[0., 1., 0, 0, 0, 0])
    rp, col, val, rhs, x0 = orc.assemble(g, t, np.zeros(24 * 25 * 26))
    x, ok, it, res, _ = orc.bicgstab2(rp, col, val, rhs, np.zeros(len(rhs)), 1e-14, 5000)
    assert ok
    i = np.arange(20 * 21 * 22) // (21 * 22)
    assert np.max(np.abs(x - (i + 0.5) / 20)) < 1e-8


def test_fixed_faces_staggered_linear_profile():
    """Fixed boundary cells (bc type 3, identity rows) on the x faces of a 2-D grid one cell
    longer along x, as a staggered representation is (CartesianRepresentation.inl:65-77),
    with the reference's staggered solve as the expectation
    (test_integration_Cartesian_Trilinos.cpp:410-496 SolveStaggeredX): phi = i / (n_x - 1)
    exactly, for N = 3 components; Neumann on the y faces. The fixed rows are identity rows
    (one entry, 1) with rhs = x0 = the face value, and rows next to them keep their neighbour
    in the fixed cell."""
    nx, ny, N = 21, 21, 3
    g = orc.grid([nx, ny], size=[21 / 20, 2.0], ghost=2, ncomp=N)
    t = orc.terms(dt=0.0, diff=1.0, bc_type=[3, 3, 2, 2], bc_value=[0.0, 1.0, 0, 0])
    phi = np.full((nx + 4) * (ny + 4) * N, 0.25)
    rp, col, val, rhs, x0 = orc.assemble(g, t, phi)
    i = np.repeat(np.arange(nx * ny) // ny, N)
    edge = (i == 0) | (i == nx - 1)
    rows = np.arange(len(rhs))
    assert np.all(np.diff(rp)[edge] == 1) and np.all(col[rp[:-1][edge]] == rows[edge])
    assert np.all(val[rp[:-1][edge]] == 1.0)
    assert np.array_equal(rhs[edge], np.where(i[edge] == 0, 0.0, 1.0)) and np.array_equal(x0[edge], rhs[edge])
    assert np.all(x0[~edge] == 0.25)
    # the first interior row still couples to the fixed cell (W neighbour present)
    r1 = N * ny                      # cell (1, 0), component 0
    assert (r1 - N * ny) in col[rp[r1]:rp[r1 + 1]]
    x, ok, it, res, _ = orc.bicgstab2(rp, col, val, rhs, x0, 1e-14, 5000)
    assert ok
    assert np.max(np.abs(x - i / (nx - 1))) < 1e-8


def _random_problem(rng, n=(5, 6, 7), ncomp=1, conv=True, react=False, periodic=False):
    g = orc.grid(list(n), size=[1.0, 1.2, 0.9], ghost=2, ncomp=ncomp,
                 periodic=[1, 1, 1] if periodic else None)
    npad = int(np.prod([m + 4 for m in n]))
    phi = rng.normal(size=npad * ncomp)
    D = np.exp(rng.normal(np.log(1e-2), 1.0, size=npad * ncomp))
    src = rng.uniform(size=npad * ncomp)
    vel = [rng.uniform(-1, 1, size=orc.lib().orc_num_faces(ctypes.byref(g), d)) for d in range(3)]
    R = None
    if react:
        k1 = rng.uniform(0.1, 10, size=int(np.prod(n)))
        k2 = rng.uniform(0.1, 10, size=int(np.prod(n)))
        R = np.stack([-k1, k2, k1, -k2], axis=1).ravel().copy()
    bc = [0] * 6 if periodic else [1, 2, 1, 1, 2, 1]
    t = orc.terms(dt=0.05, diff_cell=D, limiter=3 if conv else -1, vel=vel if conv else None,
                  source=src, react=R, bc_type=bc, bc_value=[0.3, 0, -0.2, 0.1, 0, 0.0])
    return g, t, phi


@pytest.mark.parametrize("case", ["diff", "conv", "react", "periodic"])
def test_bicgstab2_matches_dense_lu(case, rng):
    g, t, phi = _random_problem(rng, conv=case in ("conv", "periodic"), react=case == "react",
                                ncomp=2 if case == "react" else 1, periodic=case == "periodic")
    rp, col, val, rhs, x0 = orc.assemble(g, t, phi)
    n = len(rhs)
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    exact = np.linalg.solve(A, rhs)
    x, ok, it, res, hist = orc.bicgstab2(rp, col, val, rhs, x0, 1e-13 * np.linalg.norm(rhs), 500, 500)
    assert ok, (it, res)
    assert np.max(np.abs(x - exact)) <= 1e-10 * np.max(np.abs(exact))
    # the recursive residual it stops on is close to the true residual
    assert abs(np.linalg.norm(rhs - A @ x) - res) < 1e-9 * np.linalg.norm(rhs)
    assert len(hist) == it and np.all(np.isfinite(hist))


def test_bicgstab2_runs_at_least_two_passes_and_zero_rhs():
    A, b, x0 = laplace_pseudo2d()
    rp, col, val = csr_from_dense(A)
    # easy tolerance: the reference still runs two passes (k > 1 in the stop test)
    x, ok, it, res, _ = orc.bicgstab2(rp, col, val, b, x0, 1e3, 100)
    assert ok and it == 0                      # ||r0|| < tol: immediate return
    x, ok, it, res, _ = orc.bicgstab2(rp, col, val, b, x0, 0.9 * np.linalg.norm(b - A @ x0), 100)
    assert ok and it == 2
    x, ok, it, res, _ = orc.bicgstab2(rp, col, val, np.zeros_like(b), np.zeros_like(b), 1e-12, 100)
    assert ok and it == 0 and np.all(x == 0)
    x, ok, it, res, _ = orc.bicgstab2(rp, col, val, b, x0, 1e-30, 3)
    assert not ok and it == 4                  # k > max_iteration leaves k = max + 1


# ------------------------------------------------------------------ verification

def test_diffusion_n0d1_verification_1d():
    """counter_diffusion1D.cpp pattern made an assertion: Neumann at 0, Dirichlet 1 at 1."""
    n, dt, steps = 40, 1e-4, 1000
    g = orc.grid([n], ghost=2)
    t = orc.terms(dt=dt, diff=1.0, bc_type=[2, 1], bc_value=[0.0, 1.0])
    phi = np.zeros(n + 4)
    for s in range(steps):
        rp, col, val, rhs, x0 = orc.assemble(g, t, phi)
        x, ok, it, res, _ = orc.bicgstab2(rp, col, val, rhs, x0, 1e-13, 500)
        assert ok
        phi = orc.copy_to_field(g, x, phi)
    zeta = (np.arange(n) + 0.5) / n
    exact = np.array([orc.lib().orc_diffusion_n0d1(steps * dt, z, 100) for z in zeta])
    err = np.max(np.abs(phi[2:-2] - exact))
    assert err < 3e-3, err          # O(dt) + O(dx^2) discretisation error


def test_gaussian_blob_heat_kernel_c1_small():
    """C1 pattern at 24^3: implicit Euler diffusion of a Gaussian blob vs the free-space heat
    kernel (sigma^2 -> sigma^2 + 2 D t, amplitude (sigma^2/(sigma^2+2Dt))^{3/2})."""
    n, D, dt, steps, sig = 24, 1e-3, 1e-2, 10, 0.1
    g = orc.grid([n, n, n], ghost=2)
    xc = orc.cell_centers(g)
    r2 = np.sum((xc - 0.5) ** 2, axis=-1)
    phi = np.exp(-r2 / (2 * sig ** 2)).ravel()
    t = orc.terms(dt=dt, diff=D, bc_type=[1] * 6, bc_value=[0.0] * 6)
    for s in range(steps):
        rp, col, val, rhs, x0 = orc.assemble(g, t, phi)
        x, ok, it, res, _ = orc.bicgstab2(rp, col, val, rhs, x0, 1e-10 * np.linalg.norm(rhs), 200)
        assert ok
        phi = orc.copy_to_field(g, x, phi)
    T = steps * dt
    s2 = sig ** 2 + 2 * D * T
    exact = (sig ** 2 / s2) ** 1.5 * np.exp(-r2 / (2 * s2))
    inner = tuple(slice(2, -2) for _ in range(3))
    err = np.max(np.abs(phi.reshape(r2.shape)[inner] - exact[inner]))
    assert err < 2e-3, err


def _offdiag_bits(rp, col, val):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = rows != col
    return {(int(r), int(c)): int(v) for r, c, v in zip(rows[off], col[off], val[off].view(np.int64))}


@pytest.mark.parametrize("case", [
    dict(n=[6, 7, 8], periodic=None, bc=[1] * 6, dt=1e-2, size=[1.0, 1.0, 1.0]),
    dict(n=[5, 9, 12], periodic=None, bc=[1, 2, 2, 1, 1, 2], dt=5e-3, size=[0.5, 0.9, 1.7]),
    dict(n=[6, 7, 12], periodic=[1, 1, 1], bc=None, dt=5e-3, size=[1.0, 1.0, 1.0]),
    dict(n=[4, 6, 8], periodic=[1, 0, 1], bc=[0, 0, 1, 2, 0, 0], dt=0.0, size=[2.0, 1.0, 0.3]),
])
def test_constant_diffusion_assembly_is_exactly_symmetric(case):
    """One component, constant diffusivity, no convection, reaction or fixed faces: every
    off-diagonal entry equals its transpose bit for bit, the property under which the device
    fill lets the mixed solve's FP32 operator read a row's E, N, T coefficients from its
    neighbours' W, S, B. Convection breaks it. (Per-cell harmonic-mean diffusivity keeps it on
    some grids, not across a periodic wrap, so the device fill does not assume it.)"""
    rng = np.random.default_rng(11)
    go = orc.grid(case["n"], size=case["size"], ghost=2, periodic=case["periodic"])
    npad = orc.lib().orc_num_cells_padded(ctypes.byref(go))
    phi = rng.normal(size=npad)
    kw = dict(dt=case["dt"], diff=1.7e-3)
    if case["bc"] is not None:
        kw.update(bc_type=case["bc"], bc_value=list(rng.normal(size=6)))
    rp, col, val, _, _ = orc.assemble(go, orc.terms(**kw), phi)
    a = _offdiag_bits(rp, col, val)
    assert a == {(c, r): v for (r, c), v in a.items()}
    vel = [rng.uniform(-1, 1, size=orc.lib().orc_num_faces(ctypes.byref(go), d)) for d in range(3)]
    rp, col, val, _, _ = orc.assemble(go, orc.terms(limiter=4, vel=vel, **kw), phi)
    a = _offdiag_bits(rp, col, val)
    assert a != {(c, r): v for (r, c), v in a.items()}
