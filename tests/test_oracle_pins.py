"""Pins for the oracle functions that no known-answer test of the reference covers.

Each check compares the oracle (oracle/oracle.cpp) with a result the mathematics fixes:

  harmonic-mean face diffusivity  steady 1-D two-point flux through log-normal cells: the
                                  discrete solution is the series-conductance profile
  2x2 reaction block              with no diffusion, every cell is one exact implicit-Euler
                                  step (I - dt R)^-1 phi; Neumann walls conserve phi0 + phi1
  source term                     phi^{n+1} = phi^n + dt s per cell (no diffusion); Neumann
                                  walls conserve the volume integral plus dt V s
  explicit TVD step               conservation on a periodic box with random face velocity;
                                  the UPWIND step against a numpy closed form
  multi-D TVD face values         a quadratic field along each axis: r has a closed form, and
                                  the face value follows from the reference's own expression
  implicit TVD convection         A phi - rhs (the functor's convection part at phi = phi_old)
                                  equals the divergence of the explicit TVD fluxes
  convection sign                 steady 1-D upwind advection-diffusion: the discrete solution
                                  is a + b rho^i, rho = 1 + Pe (u > 0) or 1 / (1 - Pe) (u < 0)
  ghost modes 2-4                 brute force over every padded cell, plus the face properties
                                  (Dirichlet: mean of ghost and cell = v; Neumann: zero jump)
  divergence time schemes         coefficients scaled by w0 (exactly half under Crank-Nicolson),
                                  the TVD right-hand side by 1 + sum of the explicit weights

test_pins_catch_planted_mistakes compiles mutants of oracle.cpp, each with one plausible
mistake planted (on the CPU), and requires the pin aimed at it to fail.
"""
import concurrent.futures
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIMS = [0, 1, 2, 3, 4]
TOL = 1e-11           # a pin passes below this error


def dense(rp, col, val):
    n = len(rp) - 1
    A = np.zeros((n, n))
    for r in range(n):
        np.add.at(A[r], col[rp[r]:rp[r + 1]], val[rp[r]:rp[r + 1]])
    return A


def solve(g, t, phi):
    rp, col, val, rhs, x0 = orc.assemble(g, t, phi)
    return np.linalg.solve(dense(rp, col, val), rhs)


def padded_shape(g):
    return tuple(g.n[d] + 2 * g.ghost for d in range(g.dim))


def internal(g, a):
    """Internal block of a padded field (component axis last)."""
    s = tuple(slice(g.ghost, g.ghost + g.n[d]) for d in range(g.dim))
    return a.reshape(padded_shape(g) + (g.ncomp,))[s]


def nfaces(g, d):
    return orc.lib().orc_num_faces(ctypes.byref(g), d)


# ----------------------------------------------------------------------------- the pins
def pin_harmonic_series():
    """Steady -div(D grad phi) = 0 along one axis with Dirichlet vL, vR and log-normal D per
    cell (ghost D = boundary cell D): phi at cell centre i is vL + (vR - vL) R_i / R, with
    R_i = sum_{j<i} dx/D_j + dx/(2 D_i), R = sum_j dx/D_j (series conductances). Checked along
    each axis of a 3-D grid (Neumann on the other faces)."""
    rng = np.random.default_rng(5)
    err = 0.0
    for axis in range(3):
        n = [3, 3, 3]
        n[axis] = 14
        size = [0.7, 1.1, 0.9]
        g = orc.grid(n, size=size, ghost=2)
        Dl = np.exp(rng.normal(np.log(1e-2), 1.0, size=n[axis]))
        idx = np.clip(np.arange(n[axis] + 4) - 2, 0, n[axis] - 1)       # ghosts copy the boundary cell
        shape = padded_shape(g)
        Dfield = np.broadcast_to(Dl[idx].reshape([-1 if d == axis else 1 for d in range(3)]), shape).ravel().copy()
        vL, vR = 0.3, -1.7
        bc = [2] * 6
        bc[2 * axis] = bc[2 * axis + 1] = 1
        bv = [0.0] * 6
        bv[2 * axis], bv[2 * axis + 1] = vL, vR
        t = orc.terms(dt=0.0, diff_cell=Dfield, bc_type=bc, bc_value=bv)
        x = solve(g, t, np.zeros(int(np.prod(shape)))).reshape(n)
        dx = size[axis] / n[axis]
        Rc = np.cumsum(np.concatenate([[0.0], dx / Dl[:-1]])) + dx / (2 * Dl)
        exp = vL + (vR - vL) * Rc / np.sum(dx / Dl)
        got = np.moveaxis(x, axis, 0).reshape(n[axis], -1)
        err = max(err, float(np.max(np.abs(got - exp[:, None]))) / np.max(np.abs(exp)))
    return err


def pin_reaction_block():
    """No diffusion: the operator is block diagonal, so every cell holds (I - dt R_c)^-1 phi_c
    with its own R_c = [[-k1, k2], [k1, -k2]]. With diffusion and Neumann walls phi0 + phi1
    summed over the cells is conserved (the columns of R sum to zero)."""
    rng = np.random.default_rng(6)
    g = orc.grid([4, 5, 3], ghost=1, ncomp=2)
    nc = 60
    k1, k2 = rng.uniform(0.1, 10, nc), rng.uniform(0.1, 10, nc)
    R = np.stack([-k1, k2, k1, -k2], 1).ravel().copy()
    phi = rng.uniform(size=int(np.prod(padded_shape(g))) * 2)
    dt = 0.07
    x = solve(g, orc.terms(dt=dt, diff_comp=[0.0, 0.0], react=R, bc_type=[2] * 6), phi).reshape(nc, 2)
    p = internal(g, phi).reshape(nc, 2)
    err = 0.0
    for c in range(nc):
        M = np.eye(2) - dt * R[4 * c:4 * c + 4].reshape(2, 2)
        e = np.linalg.solve(M, p[c])
        err = max(err, float(np.max(np.abs(x[c] - e)) / np.max(np.abs(e))))
    x = solve(g, orc.terms(dt=dt, diff_comp=[3e-2, 7e-2], react=R, bc_type=[2] * 6), phi).reshape(nc, 2)
    err = max(err, abs(x.sum() - p.sum()) / p.sum())
    return err


def pin_source():
    """No diffusion: phi^{n+1} = phi^n + dt s cell by cell (s read at the cell's padded index).
    With diffusion and Neumann walls the volume integral grows by exactly dt V sum(s)."""
    rng = np.random.default_rng(7)
    g = orc.grid([5, 4, 6], size=[1.0, 0.8, 1.3], ghost=2)
    npad = int(np.prod(padded_shape(g)))
    phi, s = rng.uniform(size=npad), rng.uniform(size=npad)
    dt = 0.05
    x = solve(g, orc.terms(dt=dt, diff=0.0, source=s, bc_type=[2] * 6), phi)
    e = (internal(g, phi) + dt * internal(g, s)).ravel()
    err = float(np.max(np.abs(x - e)) / np.max(np.abs(e)))
    x = solve(g, orc.terms(dt=dt, diff=0.3, source=s, bc_type=[2] * 6), phi)
    tot = internal(g, phi).sum() + dt * internal(g, s).sum()
    return max(err, abs(x.sum() - tot) / tot)


def periodic_velocity(rng, g):
    """Random face velocities whose first and last face along each axis (the same face of the
    periodic box) carry the same value."""
    vel = []
    for d in range(g.dim):
        shp = [g.n[e] + (1 if e == d else 0) for e in range(g.dim)]
        u = rng.uniform(-1, 1, size=shp)
        sl_last = [slice(None)] * g.dim
        sl_first = [slice(None)] * g.dim
        sl_last[d], sl_first[d] = -1, 0
        u[tuple(sl_last)] = u[tuple(sl_first)]
        vel.append(u.ravel().copy())
    return vel


def pin_tvd_conservation():
    """Explicit TVD step on a fully periodic box (ghosts filled periodically) with random,
    non-constant face velocity: sum V phi is conserved, for every limiter, in 2-D and 3-D."""
    rng = np.random.default_rng(8)
    err = 0.0
    for n in ([6, 7, 8], [9, 5]):
        g = orc.grid(n, ghost=2, periodic=[1] * len(n))
        phi = orc.ghost_fill(g, rng.normal(size=int(np.prod(padded_shape(g)))), [1] * (2 * len(n)))
        vel = periodic_velocity(rng, g)
        for lim in LIMS:
            out = orc.tvd_explicit_step(g, lim, phi, vel, 0.01)
            a, b = internal(g, phi).sum(), internal(g, out).sum()
            err = max(err, abs(b - a) / np.abs(internal(g, phi)).sum())
    return err


def pin_tvd_quadratic():
    """phi = m^2 along axis d (m: index relative to the first internal cell, ghosts included)
    plus 7 x the index along the next axis. Upwind cell U, its upwind neighbour UU and the
    downwind cell D of face m (between cells m-1 and m) are m-1, m-2, m for u >= 0 and m, m+1,
    m-1 for u < 0, so r = (2m-3)/(2m-1) or (2m+1)/(2m-1) with every difference exact. The face
    flux must equal u times the reference's face expression of those three values."""
    R = orc.ref()
    rng = np.random.default_rng(9)
    err = 0.0
    g = orc.grid([6, 5, 7], ghost=2)
    shape = padded_shape(g)
    for d in range(3):
        m = np.arange(shape[d]) - 2
        e = (d + 1) % 3
        me = np.arange(shape[e]) - 2
        phi = (np.reshape(m.astype(float) ** 2, [-1 if a == d else 1 for a in range(3)])
               + 7.0 * np.reshape(me.astype(float), [-1 if a == e else 1 for a in range(3)]))
        phi = np.broadcast_to(phi, shape).ravel().copy()
        vel = [rng.uniform(-1, 1, size=nfaces(g, a)) for a in range(3)]
        for lim in LIMS:
            f = orc.tvd_face_flux(g, lim, phi, vel)[d]
            fshape = [g.n[a] + (1 if a == d else 0) for a in range(3)]
            u = vel[d].reshape(fshape)
            F = f.reshape(fshape)
            P = phi.reshape(shape)
            for q in np.ndindex(*fshape):
                mm = q[d]
                cells = (mm - 1, mm - 2, mm) if u[q] >= 0 else (mm, mm + 1, mm - 1)   # U, UU, D
                vals = []
                for c in cells:
                    p = [q[a] + 2 for a in range(3)]
                    p[d] = c + 2
                    vals.append(P[tuple(p)])
                pU, pUU, pD = vals
                if lim == 4:     # van Leer (not in the reference): its closed form
                    den = pD - pU
                    rr = (pU - pUU) / den if den != 0 else np.nan
                    psi = orc.lib().orc_limiter(4, rr)
                    face = pU + (psi * 0.5) * (pD - pU)
                else:
                    face = R.ref_tvd_face_value(lim, pUU, pU, pD)
                exp = face * u[q]
                err = max(err, abs(F[q] - exp) / (1.0 + abs(exp)))
    return err


def pin_tvd_upwind_step():
    """UPWIND explicit step against its numpy closed form (vectorised, per axis): face flux
    u * phi(upwind cell), phi_out = phi - dt/V sum_d A_d (F_hi - F_lo)."""
    rng = np.random.default_rng(10)
    g = orc.grid([7, 6, 5], size=[1.0, 0.6, 1.4], ghost=2)
    shape = padded_shape(g)
    P = rng.normal(size=shape)
    vel = [rng.uniform(-1, 1, size=nfaces(g, d)) for d in range(3)]
    dt = 0.013
    out = orc.tvd_explicit_step(g, 1, P.ravel().copy(), vel, dt)
    dx = [g.size[d] / g.n[d] for d in range(3)]
    V = dx[0] * dx[1] * dx[2]
    inner = P[2:-2, 2:-2, 2:-2]
    div = np.zeros_like(inner)
    for d in range(3):
        A = V / dx[d]
        fshape = [g.n[a] + (1 if a == d else 0) for a in range(3)]
        u = vel[d].reshape(fshape)
        lo_sl = [slice(2, -2)] * 3
        hi_sl = [slice(2, -2)] * 3
        lo_sl[d] = slice(1, 2 + g.n[d])          # cell left of face m: m - 1
        hi_sl[d] = slice(2, 3 + g.n[d])          # cell right of face m: m
        left, right = P[tuple(lo_sl)], P[tuple(hi_sl)]
        F = u * np.where(u >= 0, left, right)
        f_hi = [slice(None)] * 3
        f_lo = [slice(None)] * 3
        f_hi[d], f_lo[d] = slice(1, None), slice(0, -1)
        div += A * (F[tuple(f_hi)] - F[tuple(f_lo)])
    exp = inner - dt / V * div
    got = out.reshape(shape)[2:-2, 2:-2, 2:-2]
    return float(np.max(np.abs(got - exp)) / np.max(np.abs(exp)))


def pin_tvd_matrix_identity():
    """At phi = phi_old the functor's implicit convection reproduces the explicit flux
    divergence: (A phi)_i - rhs_i = sum_d A_d (F_hi - F_lo) with F the TVD face fluxes, since
    the upwind part carries u phi_U and the deferred correction u (phi_f - phi_U). Periodic
    2-D and 3-D boxes, random non-constant face velocity, every limiter."""
    rng = np.random.default_rng(11)
    err = 0.0
    for n in ([5, 6, 7], [8, 6]):
        g = orc.grid(n, size=[1.0, 0.9, 1.2][:len(n)], ghost=2, periodic=[1] * len(n))
        phi = orc.ghost_fill(g, rng.normal(size=int(np.prod(padded_shape(g)))), [1] * (2 * len(n)))
        vel = periodic_velocity(rng, g)
        for lim in LIMS:
            rp, col, val, rhs, x0 = orc.assemble(g, orc.terms(dt=0.0, diff=0.0, limiter=lim, vel=vel), phi)
            lhs = orc.spmv(rp, col, val, internal(g, phi).ravel()) - rhs
            div = internal(g, orc.div_faces(g, orc.tvd_face_flux(g, lim, phi, vel))).ravel()
            err = max(err, float(np.max(np.abs(lhs - div)) / np.max(np.abs(div))))
    return err


def pin_convection_sign():
    """Steady upwind advection-diffusion along one axis (dt = 0, UPWIND, constant u, Dirichlet
    vL, vR, Neumann elsewhere). Interior rows read -D/dx (phi_{i+1} - 2 phi_i + phi_{i-1}) +
    u (phi_i - phi_{i-1}) = 0 (u > 0; u (phi_{i+1} - phi_i) for u < 0), so phi_i = a + b rho^i
    with rho = 1 + Pe (u > 0) or 1 / (1 - Pe) (u < 0), Pe = u dx / D, also through the mirror
    ghosts phi_{-1}, phi_n whose means with phi_0, phi_{n-1} are vL, vR."""
    err = 0.0
    for axis, u in ((0, 0.8), (1, -0.6), (2, 1.3)):
        n = [2, 2, 2]
        n[axis] = 20
        size = [1.0, 1.0, 1.0]
        g = orc.grid(n, size=size, ghost=2)
        D = 0.04
        vel = []
        for d in range(3):
            vel.append(np.full(nfaces(g, d), u if d == axis else 0.0))
        bc = [2] * 6
        bc[2 * axis] = bc[2 * axis + 1] = 1
        vL, vR = 0.3, 1.1
        bv = [0.0] * 6
        bv[2 * axis], bv[2 * axis + 1] = vL, vR
        t = orc.terms(dt=0.0, diff=D, limiter=1, vel=vel, bc_type=bc, bc_value=bv)
        x = solve(g, t, np.zeros(int(np.prod(padded_shape(g))))).reshape(n)
        dx = size[axis] / n[axis]
        Pe = u * dx / D
        rho = 1 + Pe if u > 0 else 1 / (1 - Pe)
        N = n[axis]
        M = np.array([[1.0, (rho ** -1 + 1) / 2], [1.0, (rho ** (N - 1) + rho ** N) / 2]])
        a, b = np.linalg.solve(M, [vL, vR])
        exp = a + b * rho ** np.arange(N)
        got = np.moveaxis(x, axis, 0).reshape(N, -1)
        err = max(err, float(np.max(np.abs(got - exp[:, None])) / np.max(np.abs(exp))))
    return err


def ghost_brute(g, phi, modes, values):
    """Ghost fill by brute force: a padded cell outside the internal range along exactly one
    axis d, in layer l (1..g) of face f, takes phi of the mirror cell (layer l inside the face)
    for modes 2 (2 v - phi) and 3 (phi), v for mode 4, the wrapped cell for mode 1."""
    shape = padded_shape(g)
    P = phi.reshape(shape + (g.ncomp,)).copy()
    src = phi.reshape(shape + (g.ncomp,))
    for idx in np.ndindex(*shape):
        out = [d for d in range(g.dim) if not (g.ghost <= idx[d] < g.ghost + g.n[d])]
        if len(out) != 1:
            continue
        d = out[0]
        i = idx[d] - g.ghost
        low = i < 0
        f = 2 * d + (0 if low else 1)
        mode = modes[f]
        if mode == 0:
            continue
        m = list(idx)
        if mode == 1:
            m[d] = i % g.n[d] + g.ghost
        else:
            l = -i if low else i - g.n[d] + 1
            m[d] = g.ghost + (l - 1 if low else g.n[d] - l)
        s = src[tuple(m)]
        P[idx] = {1: s, 2: 2 * values[f] - s, 3: s, 4: np.full(g.ncomp, values[f])}[mode]
    return P.ravel()


def pin_ghost_modes():
    rng = np.random.default_rng(12)
    err = 0.0
    for n, modes, per in (([4, 5, 6], [2, 3, 4, 2, 3, 0], None), ([5, 7], [3, 2, 4, 4], None),
                          ([6, 4, 5], [1, 1, 2, 4, 3, 2], [1, 0, 0])):
        g = orc.grid(n, ghost=2, ncomp=2, periodic=per)
        phi = rng.normal(size=int(np.prod(padded_shape(g))) * 2)
        vals = list(rng.normal(size=6))
        got = orc.ghost_fill(g, phi, modes, vals)
        err = max(err, float(np.max(np.abs(got - ghost_brute(g, phi, modes, vals)))))
        # face properties: Dirichlet mean = v, Neumann zero jump (first ghost layer)
        P = got.reshape(padded_shape(g) + (2,))
        for f, mode in enumerate(modes[:2 * g.dim]):
            d, hi = f // 2, f % 2
            gi = g.ghost + g.n[d] if hi else g.ghost - 1
            ci = g.ghost + g.n[d] - 1 if hi else g.ghost
            sl = [slice(g.ghost, g.ghost + g.n[e]) for e in range(g.dim)]
            sg, sc = list(sl), list(sl)
            sg[d], sc[d] = gi, ci
            if mode == 2:
                err = max(err, float(np.max(np.abs(0.5 * (P[tuple(sg)] + P[tuple(sc)]) - vals[f]))))
            elif mode == 3:
                err = max(err, float(np.max(np.abs(P[tuple(sg)] - P[tuple(sc)]))))
    return err


def pin_time_scheme():
    """Divergence<Grid, TimeScheme> in matrix form reuses the implicit level's face stencil for
    the explicit levels (Divergence_Cartesian.inl:58-66,183-204). Against the same operator
    under EULER_BACKWARD: every diffusion and convection coefficient is scaled by w0
    (CRANK_NICHOLSON's off-diagonals are exactly half of EULER_BACKWARD's, 0.5 being a power of
    two; EULER_FORWARD's and ADAMS_BASHFORT's vanish), and on a periodic box the convection's
    right-hand side (the TVD deferred correction) is scaled by 1 + sum_{t>=1} w_t (x1.5 for
    CRANK_NICHOLSON, x2 for EULER_FORWARD and ADAMS_BASHFORT, x19/12 for ADAMS_MOULTON) while
    the DDT part is untouched. Returns the largest relative error."""
    rng = np.random.default_rng(41)
    schemes = {"CN": (0.5, 0.5), "EF": (0.0, 1.0), "AB": (0.0, 1.5, -0.5), "AM": (5 / 12, 8 / 12, -1 / 12)}
    err = 0.0
    # diffusion (per-cell diffusivity, Dirichlet / Neumann faces), no DDT
    g = orc.grid([6, 5, 7], size=[1.0, 0.8, 1.2], ghost=2)
    npad = int(np.prod(padded_shape(g)))
    phi = rng.normal(size=npad)
    D = np.exp(rng.normal(np.log(1e-3), 1.0, size=npad))
    kw = dict(diff_cell=D, bc_type=[1, 2, 1, 1, 2, 1], bc_value=list(rng.normal(size=6)))
    Ae = dense(*orc.assemble(g, orc.terms(**kw), phi)[:3])
    off = ~np.eye(len(Ae), dtype=bool)
    scale = float(np.max(np.abs(Ae[off])))
    for name, w in schemes.items():
        As = dense(*orc.assemble(g, orc.terms(div_scheme=w, **kw), phi)[:3])
        if name == "CN":      # exact: a power-of-two scaling
            err = max(err, float(np.max(np.abs(As[off] - 0.5 * Ae[off]))) / scale * 1e6)
        err = max(err, float(np.max(np.abs(As[off] - w[0] * Ae[off]))) / scale)
    # convection (VANLEER, random face velocities) + DDT on a periodic box
    gp = orc.grid([6, 5, 7], size=[1.0, 0.8, 1.2], ghost=2, periodic=[1, 1, 1])
    npad = int(np.prod(padded_shape(gp)))
    phi = rng.normal(size=npad)
    vel = [rng.uniform(-1, 1, size=nfaces(gp, d)) for d in range(3)]
    ddt = orc.assemble(gp, orc.terms(dt=1e-2), phi)
    e = orc.assemble(gp, orc.terms(dt=1e-2, limiter=4, vel=vel), phi)
    conv_e = e[3] - ddt[3]
    Ae = dense(*e[:3]) - dense(*ddt[:3])
    off = ~np.eye(len(Ae), dtype=bool)
    rscale, cscale = float(np.max(np.abs(conv_e))), float(np.max(np.abs(Ae[off])))
    for name, w in schemes.items():
        s_ = orc.assemble(gp, orc.terms(dt=1e-2, limiter=4, vel=vel, div_scheme=w), phi)
        factor = 1.0 + sum(w[1:])
        err = max(err, float(np.max(np.abs((s_[3] - ddt[3]) - factor * conv_e))) / rscale)
        As = dense(*s_[:3]) - dense(*ddt[:3])
        err = max(err, float(np.max(np.abs(As[off] - w[0] * Ae[off]))) / cscale)
    return err


PINS = {
    "harmonic_series": pin_harmonic_series,
    "reaction_block": pin_reaction_block,
    "source": pin_source,
    "tvd_conservation": pin_tvd_conservation,
    "tvd_quadratic": pin_tvd_quadratic,
    "tvd_upwind_step": pin_tvd_upwind_step,
    "tvd_matrix_identity": pin_tvd_matrix_identity,
    "convection_sign": pin_convection_sign,
    "ghost_modes": pin_ghost_modes,
    "time_scheme": pin_time_scheme,
}


def _need_ref(name):
    if name == "tvd_quadratic" and orc.ref() is None:
        pytest.skip("oracle/_ref not built")


@pytest.mark.parametrize("name", list(PINS))
def test_oracle_pin(name):
    _need_ref(name)
    err = PINS[name]()
    assert err < TOL, (name, err)


def test_constant_diff_cell_equals_constant_diffusivity():
    """A uniform per-cell diffusivity reproduces the constant-diffusivity assembly bit for bit
    (the harmonic mean of equal values D is D for this D)."""
    D = 1e-3
    assert ((2.0 * D) * D) / (D + D) == D
    rng = np.random.default_rng(13)
    g = orc.grid([5, 6, 7], size=[1.0, 0.7, 1.3], ghost=2)
    npad = int(np.prod(padded_shape(g)))
    phi = rng.normal(size=npad)
    kw = dict(dt=0.02, bc_type=[1, 2, 1, 1, 2, 1], bc_value=list(rng.normal(size=6)))
    a = orc.assemble(g, orc.terms(diff=D, **kw), phi)
    b = orc.assemble(g, orc.terms(diff_cell=np.full(npad, D), **kw), phi)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


def test_per_cell_diffusivity_operator_is_exactly_symmetric():
    """DDT + per-cell diffusivity (harmonic face means) + source with Dirichlet / Neumann
    faces and no periodic axis: A == A^T bit for bit (the harmonic mean takes its two cell
    values in face order from either side), the property the device solver's four-slot
    traversal relies on. Across a periodic wrap the face reads ghost diffusivities, and with
    ghosts that are not the periodic images the operator is not symmetric."""
    import scipy.sparse as sp

    def bitwise_symmetric(rp, col, val):
        n = len(rp) - 1
        M = sp.csr_matrix((val, col, rp), shape=(n, n))
        T = M.T.tocsr()
        M.sort_indices()
        T.sort_indices()
        return (np.array_equal(M.indices, T.indices) and np.array_equal(M.indptr, T.indptr)
                and np.array_equal(M.data.view(np.uint64), T.data.view(np.uint64)))

    rng = np.random.default_rng(29)
    for n, size, bc in [([6, 7, 8], [1.0, 0.8, 1.1], [1, 2, 1, 1, 2, 1]), ([5, 9, 4], [0.5, 1.0, 0.7], [1] * 6)]:
        g = orc.grid(n, size=size, ghost=2)
        npad = int(np.prod(padded_shape(g)))
        D = np.exp(rng.normal(np.log(1e-3), 1.0, size=npad))
        t = orc.terms(dt=5e-3, diff_cell=D, source=rng.uniform(size=npad), bc_type=bc,
                      bc_value=list(rng.normal(size=6)))
        rp, col, val, _, _ = orc.assemble(g, t, rng.normal(size=npad))
        assert bitwise_symmetric(rp, col, val)
    g = orc.grid([6, 7, 8], ghost=2, periodic=[1, 1, 1])
    npad = int(np.prod(padded_shape(g)))
    t = orc.terms(dt=5e-3, diff_cell=np.exp(rng.normal(np.log(1e-3), 1.0, size=npad)))
    rp, col, val, _, _ = orc.assemble(g, t, rng.normal(size=npad))
    assert not bitwise_symmetric(rp, col, val)


# ----------------------------------------------------------------------------- planted mistakes
# (mutant name, text in oracle.cpp, replacement, pins that must fail)
MUTANTS = [
    ("arithmetic_mean", "Df = ((2. * Dlo) * Dhi) / (Dlo + Dhi);", "Df = 0.5 * (Dlo + Dhi);", ["harmonic_series"]),
    ("reaction_sign", "mb->get(c, 0) -= G.V * R[c * 2 + c];", "mb->get(c, 0) += G.V * R[c * 2 + c];",
     ["reaction_block"]),
    ("reaction_transposed", "-(G.V * R[c * 2 + o])", "-(G.V * R[o * 2 + c])", ["reaction_block"]),
    ("source_volume", "mb->rhs[c] += G.V * t->source[lp * G.N + c];", "mb->rhs[c] += t->source[lp * G.N + c];",
     ["source"]),
    ("swapped_face_velocity", "uf[2 * d] = vel[d][face_index(G, d, q)];\n        uf[2 * d + 1] = vel[d][face_index(G, d, qh)];",
     "uf[2 * d] = vel[d][face_index(G, d, qh)];\n        uf[2 * d + 1] = vel[d][face_index(G, d, q)];",
     ["tvd_conservation", "tvd_upwind_step"]),
    ("upwind_upwind_cell", "UU->v[d] += vn - 2 * vp;", "UU->v[d] += vn - vp;", ["tvd_quadratic"]),
    ("explicit_step_sign", "phi_out[lp * G.N + c] = phi[lp * G.N + c] - fac * div_value(G, s, c);",
     "phi_out[lp * G.N + c] = phi[lp * G.N + c] + fac * div_value(G, s, c);", ["tvd_upwind_step"]),
    ("convection_subtracted", "mb->add(div_matrix(G, tvd_matrix(G, t->conv_limiter, phi, p, uf), dw, dn));",
     "mb->subtract(div_matrix(G, tvd_matrix(G, t->conv_limiter, phi, p, uf), dw, dn));",
     ["convection_sign", "tvd_matrix_identity"]),
    ("correction_sign", "s.rhs[c][f] = ((psi * -0.5) * (pD - pU)) * vel;", "s.rhs[c][f] = ((psi * 0.5) * (pD - pU)) * vel;",
     ["tvd_matrix_identity"]),
    ("ghost_mirror_layer", "else src.v[d] = G.g + (low ? l - 1 : G.n[d] - l);",
     "else src.v[d] = G.g + (low ? l : G.n[d] - l - 1);", ["ghost_modes"]),
    ("dirichlet_ghost", "case 2: *dst = 2. * value[f] - s; break;", "case 2: *dst = value[f] - s; break;",
     ["ghost_modes"]),
    ("scheme_level_unweighted", "o.rhs[c] += w[t] * G.A[0] * (s.rhs[c][1] - s.rhs[c][0]);",
     "o.rhs[c] += G.A[0] * (s.rhs[c][1] - s.rhs[c][0]);", ["time_scheme"]),
    ("scheme_coefficients_unscaled", "for (int q = 0; q < 7; q++) o.a[c][q] *= w0;",
     "for (int q = 0; q < 7; q++) o.a[c][q] *= 1.;", ["time_scheme"]),
]


def _build_mutant(tmp, name, old, new):
    src = open(os.path.join(ROOT, "oracle", "oracle.cpp")).read()
    assert src.count(old) == 1, f"mutation target of {name} not found exactly once in oracle.cpp"
    path = os.path.join(tmp, f"{name}.cpp")
    with open(path, "w") as f:
        f.write(src.replace(old, new))
    out = os.path.join(tmp, f"lib_{name}.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                    "-I", os.path.join(ROOT, "oracle"), "-o", out, path], check=True, capture_output=True)
    return out


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    tmp = str(tmp_path_factory.mktemp("mutants"))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = {m[0]: ex.submit(_build_mutant, tmp, m[0], m[1], m[2]) for m in MUTANTS}
        return {k: f.result() for k, f in futs.items()}


@pytest.mark.parametrize("mutant", [m[0] for m in MUTANTS])
def test_pins_catch_planted_mistakes(mutant, mutant_libs):
    pins = next(m[3] for m in MUTANTS if m[0] == mutant)
    with orc.using(mutant_libs[mutant]):
        for name in pins:
            _need_ref(name)
            try:
                err = PINS[name]()
            except (AssertionError, np.linalg.LinAlgError):
                err = np.inf          # the mutant broke the oracle's own consistency checks
            assert err > 1e3 * TOL, (mutant, name, err)


# ----------------------------------------------------------------------------- face interpolation
def test_interpolate_to_scalar_face_kat():
    """InterpolateToScalarFaceTest (Grid/test/test_integration_Cartesian_Interpolation.cpp:
    1075-1260) for a cell-centred source and target: at cell (2, 3, 4) the WEST face value is
    0.5 (f(ind) + f(ind - e_i)) and the TOP face value 0.5 (f(ind) + f(ind + e_k)), within
    1e3 eps, with f drawn as 1e-4 x an integer in [-10000, 10000]. Every other face too, and
    bit for bit against the reference's order (neighbour + cell) x Divisor<2>."""
    rng = np.random.default_rng(14)
    g = orc.grid([5, 6, 7], ghost=2)
    shape = padded_shape(g)
    u = [1e-4 * rng.integers(-10000, 10001, size=int(np.prod(shape))).astype(float) for _ in range(3)]
    f = orc.interpolate_to_faces(g, u)
    eps = np.finfo(float).eps
    ind = (2, 3, 4)                                      # internal index
    for d, face_hi in ((0, False), (2, True)):           # WEST, TOP
        U = u[d].reshape(shape)
        p = tuple(i + 2 for i in ind)
        nb = list(p)
        nb[d] += 1 if face_hi else -1
        v_ex = 0.5 * (U[p] + U[tuple(nb)])
        q = list(ind)
        if face_hi:
            q[d] += 1
        fshape = [g.n[a] + (1 if a == d else 0) for a in range(3)]
        v = f[d].reshape(fshape)[tuple(q)]
        assert abs(v - v_ex) <= 1e3 * eps * abs(v_ex)
    for d in range(3):
        U = u[d].reshape(shape)
        fshape = [g.n[a] + (1 if a == d else 0) for a in range(3)]
        F = f[d].reshape(fshape)
        for q in np.ndindex(*fshape):
            hi = tuple(q[a] + 2 for a in range(3))
            lo = list(hi)
            lo[d] -= 1
            assert F[q] == (U[tuple(lo)] + U[hi]) * 0.5


def test_interpolated_linear_velocity_is_exact_at_faces():
    """A velocity linear in the coordinate along its axis is reproduced exactly (to rounding)
    at the face centres: the two-point mean of a linear function is its midpoint value."""
    g = orc.grid([6, 5, 4], size=[1.0, 0.5, 2.0], ghost=2)
    xc = orc.cell_centers(g)
    u = [(0.3 + 1.7 * xc[..., d]).ravel().copy() for d in range(3)]
    f = orc.interpolate_to_faces(g, u)
    for d in range(3):
        dx = g.size[d] / g.n[d]
        fshape = [g.n[a] + (1 if a == d else 0) for a in range(3)]
        xf = np.arange(fshape[d]) * dx
        exp = np.broadcast_to((0.3 + 1.7 * xf).reshape([-1 if a == d else 1 for a in range(3)]), fshape)
        assert np.max(np.abs(f[d].reshape(fshape) - exp)) <= 1e-14 * np.max(np.abs(exp))
