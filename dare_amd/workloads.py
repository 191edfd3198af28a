"""Synthetic inputs of the benchmark configurations (seeded, generated on the device).

c2_rotation: 256^3 solid-body rotation u = 2 pi (-(y-1/2), x-1/2, 0) on cell faces, a step
    cube (1 inside a cube of side 0.25 centred at (0.5, 0.75, 0.5)), CFL 0.4 explicit TVD
    advection, implicit diffusion D = 1e-3 with Dirichlet-0 walls.
c1_blob: 64^3 Gaussian blob (centre 0.5, sigma 0.1), D = 1e-3, dt = 1e-2, Dirichlet 0.
c3_lognormal: cell diffusivity ln D ~ N(ln 1e-3, 1), uniform source in [0,1), Dirichlet 0.
"""
import math

import numpy as np

from . import BC_DIRICHLET, GHOST_CONSTANT, Grid, Terms


def _torch():
    import torch
    return torch


def cell_centers_1d(n, size, ghost):
    dx = size / n
    return (np.arange(n + 2 * ghost) - ghost + 0.5) * dx


def face_positions_1d(n, size):
    dx = size / n
    return np.arange(n + 1) * dx


class C2Rotation:
    """Config C2 state on the device (field with ghosts, face velocities, terms)."""

    def __init__(self, n=256, device="cuda", D=1e-3, cfl=0.4):
        torch = _torch()
        self.grid = Grid.make([n, n, n], size=[1.0, 1.0, 1.0], ghost=2)
        g = self.grid
        self.n = n
        self.dx = 1.0 / n
        dt = torch.float64
        xc = torch.tensor(cell_centers_1d(n, 1.0, 2), dtype=dt, device=device)
        # face velocities: x faces (n+1, n, n) at (x_f, y_c, z_c); y faces (n, n+1, n)
        yc_int = xc[2:-2]
        ux = (-(2 * math.pi) * (yc_int - 0.5)).view(1, n, 1).expand(n + 1, n, n).contiguous()
        uy = ((2 * math.pi) * (yc_int - 0.5)).view(n, 1, 1).expand(n, n + 1, n).contiguous()
        uz = torch.zeros((n, n, n + 1), dtype=dt, device=device)
        self.vel = [ux.view(-1), uy.view(-1), uz.view(-1)]
        umax = max(float(ux.abs().max()), float(uy.abs().max()), float(uz.abs().max()))
        self.dt = cfl * self.dx / umax
        X = xc.view(-1, 1, 1)
        Y = xc.view(1, -1, 1)
        Z = xc.view(1, 1, -1)
        inside = ((X - 0.5).abs() < 0.125) & ((Y - 0.75).abs() < 0.125) & ((Z - 0.5).abs() < 0.125)
        self.phi = inside.to(dt).contiguous().view(-1)
        self.phi_work = self.phi.clone()
        self.D = D
        self.terms = Terms.make(dt=self.dt, diff=D, bc_type=[BC_DIRICHLET] * 6, bc_value=[0.0] * 6)
        self.ghost_modes = [GHOST_CONSTANT] * 6
        self.ghost_values = [0.0] * 6
        ncell = n ** 3
        self.rhs = torch.zeros(ncell, dtype=dt, device=device)
        self.x = torch.zeros(ncell, dtype=dt, device=device)


def c1_blob_field(n=64, device="cuda", sigma=0.1):
    torch = _torch()
    xc = torch.tensor(cell_centers_1d(n, 1.0, 2), dtype=torch.float64, device=device)
    r2 = (xc.view(-1, 1, 1) - 0.5) ** 2 + (xc.view(1, -1, 1) - 0.5) ** 2 + (xc.view(1, 1, -1) - 0.5) ** 2
    return torch.exp(-r2 / (2 * sigma ** 2)).contiguous().view(-1)


def c3_lognormal(n, device="cuda", seed=3, mean_diff_number=1.0):
    """Diffusivity and source of the precision form at resolution n (padded layout)."""
    torch = _torch()
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    npad = (n + 4) ** 3
    lnD = torch.randn(npad, generator=gen, dtype=torch.float64, device=device) + math.log(1e-3)
    D = torch.exp(lnD)
    src = torch.rand(npad, generator=gen, dtype=torch.float64, device=device)
    dx = 1.0 / n
    # pick dt so that the mean of D dt / dx^2 equals mean_diff_number
    dt = mean_diff_number * dx * dx / float(D.mean())
    return D, src, dt


class C4Batch:
    """Batch of independent implicit 2-component reaction-diffusion systems.

    Per system: edge lengths n0, n1, n2 log-uniform in [nmin, nmax] (isotropic cells of width
    1/64), D1, D2 ~ U[1e-4, 1e-2] (one per component), and per cell a linearised reaction
    block [[-k1, k2], [k1, -k2]] with k1, k2 ~ U[0.1, 10]; implicit Euler with dt, Dirichlet 0
    on every face; phi_old ~ U[0, 1) per component. Each system's structure is built and its
    values assembled straight into the rows of the batch's concatenated b and x.
    """

    def __init__(self, ctx, nsys=512, seed=4, nmin=16, nmax=96, dt=1e-2, device="cuda"):
        from . import Batch, Csr, Grid, Terms, build_structure, check, lib
        torch = _torch()
        rng = np.random.default_rng(seed)
        dims = np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=(nsys, 3)))
        self.dims = np.clip(np.rint(dims), nmin, nmax).astype(int)
        self.D = rng.uniform(1e-4, 1e-2, size=(nsys, 2))
        self.dt = dt
        self.nsys = nsys
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
        self.grids, self.terms, self.csr, self.phi, self.react = [], [], [], [], []
        h = 1.0 / 64
        for s in range(nsys):
            n = [int(v) for v in self.dims[s]]
            g = Grid.make(n, size=[v * h for v in n], ghost=1, ncomp=2)
            ncell = n[0] * n[1] * n[2]
            k = 0.1 + 9.9 * torch.rand((ncell, 2), generator=gen, dtype=torch.float64, device=device)
            R = torch.stack([-k[:, 0], k[:, 1], k[:, 0], -k[:, 1]], dim=1).contiguous().view(-1)
            t = Terms.make(dt=dt, diff_comp=[float(self.D[s, 0]), float(self.D[s, 1])], react=R,
                           bc_type=[BC_DIRICHLET] * 6, bc_value=[0.0] * 6)
            A = Csr(ctx)
            build_structure(ctx, A, g, t)
            npad = (n[0] + 2) * (n[1] + 2) * (n[2] + 2) * 2
            phi = torch.rand(npad, generator=gen, dtype=torch.float64, device=device)
            self.grids.append(g)
            self.terms.append(t)
            self.csr.append(A)
            self.phi.append(phi)
            self.react.append(R)
        self.batch = Batch(ctx, self.csr)
        self.ctx = ctx
        self._check, self._lib = check, lib

    @property
    def cells(self):
        return int(np.prod(self.dims, axis=1).sum())

    def assemble(self):
        """Values, rhs and x0 (= phi_old) of every system, into the batch's vectors."""
        from . import fill_values
        for s in range(self.nsys):
            fill_values(self.ctx, self.csr[s], self.grids[s], self.terms[s], self.phi[s],
                        self.batch.rhs_ptr(s), self.batch.x_ptr(s))
        self.ctx.synchronize()

    def close(self):
        self.batch.close()
        for A in self.csr:
            A.close()


def _base_guess(W, with_residual=True, out=None):
    """xb = 2 phi^n - phi^{n-1} in solver rows (x holds phi^n after the fill) into out (default
    W.xb) and, with with_residual, r0 = rhs - A xb."""
    from . import residual
    g = W.grid.ghost
    n = W.n
    xb = W.xb if out is None else out
    prev_int = W.phi_prev.view(n + 2 * g, n + 2 * g, n + 2 * g)[g:-g, g:-g, g:-g]
    # one elementwise pass over x and the interior view of phi^{n-1}: lerp(p, x, 2) = x + (x - p)
    _torch().lerp(prev_int, W.x.view(n, n, n), 2.0, out=xb.view(n, n, n))
    if with_residual:
        residual(W.ctx, W.A, W.rhs, xb, W.r0)


class _BaseGuess:
    def base_guess(self, with_residual=True, out=None):
        """the extrapolated guess xb (and its residual r0) of the assembled step, into out
        (default self.xb); callers that time a guess call this inside the timed region
        (assemble(base=False) skips it)"""
        _base_guess(self, with_residual, out)


def gaussian_sum(n, rng, count=8, ghost=2, device="cuda"):
    """sum of `count` Gaussians, centres U(unit cube), sigma U[0.05, 0.2], amplitude U[-1, 1],
    at the cell centres of the padded grid (ghosts included; callers overwrite them)."""
    torch = _torch()
    xc = torch.tensor(cell_centers_1d(n, 1.0, ghost), dtype=torch.float64, device=device)
    X, Y, Z = xc.view(-1, 1, 1), xc.view(1, -1, 1), xc.view(1, 1, -1)
    phi = torch.zeros((n + 2 * ghost,) * 3, dtype=torch.float64, device=device)
    c = rng.uniform(0.0, 1.0, size=(count, 3))
    sig = rng.uniform(0.05, 0.2, size=count)
    amp = rng.uniform(-1.0, 1.0, size=count)
    for q in range(count):
        r2 = (X - c[q, 0]) ** 2 + (Y - c[q, 1]) ** 2 + (Z - c[q, 2]) ** 2
        phi += amp[q] * torch.exp(-r2 / (2 * sig[q] ** 2))
    return phi.contiguous().view(-1)


class C5AdvDiff(_BaseGuess):
    """Transient advection-diffusion trajectory on an n^3 unit cube (learned-form workload).

    Per seed: D ~ U[1e-4, 1e-2], constant velocity u ~ U[-1, 1]^3 on every face, initial field
    a sum of 8 Gaussians; implicit Euler with dt, van Leer TVD convection (upwind implicit part
    + deferred correction from phi^n), Dirichlet 0 walls (ghost cells held at 0). One shared
    matrix structure serves every seed of a resolution (pass it as `A`).
    """

    @staticmethod
    def in_channels_for():
        return 3


    def __init__(self, ctx, seed, n=128, dt=5e-3, A=None, device="cuda"):
        from . import VANLEER, Csr, build_structure, lib as _lib
        import ctypes
        torch = _torch()
        rng = np.random.default_rng(seed)
        self.ctx, self.n, self.dt, self.seed = ctx, n, dt, seed
        self.D = float(rng.uniform(1e-4, 1e-2))
        self.u = rng.uniform(-1.0, 1.0, size=3)
        self.grid = Grid.make([n, n, n], size=[1.0, 1.0, 1.0], ghost=2)
        self.vel = []
        for d in range(3):
            nf = [n, n, n]
            nf[d] += 1
            self.vel.append(torch.full((nf[0] * nf[1] * nf[2],), float(self.u[d]), dtype=torch.float64,
                                       device=device))
        self.terms = Terms.make(dt=dt, diff=self.D, limiter=VANLEER, vel=self.vel, bc_type=[BC_DIRICHLET] * 6,
                                bc_value=[0.0] * 6)
        self.phi = gaussian_sum(n, rng, device=device)
        self.ghost_modes, self.ghost_values = [GHOST_CONSTANT] * 6, [0.0] * 6
        self.phi_prev = self.phi.clone()
        self.rhs = torch.zeros(n ** 3, dtype=torch.float64, device=device)
        self.x = torch.zeros_like(self.rhs)
        self.r0 = torch.zeros_like(self.rhs)
        self.xb = torch.zeros_like(self.rhs)
        self.A = A
        if self.A is None:
            self.A = Csr(ctx)
            build_structure(ctx, self.A, self.grid, self.terms)
        self.rhs_scale = dt / (1.0 / n) ** 3      # dt / V: residuals of the step in field units
        self.step_index = 0
        self.extras = []
        self.in_channels = 3

    def assemble(self, base=True):
        """ghost fill, then values / rhs / x0 = phi^n of the next implicit step; with base, also
        the extrapolated guess xb = 2 phi^n - phi^{n-1} and its residual r0 = rhs - A xb (the
        warm-start network's base guess and input)"""
        from . import fill_values, ghost_fill
        ghost_fill(self.ctx, self.grid, self.phi, self.ghost_modes, self.ghost_values)
        fill_values(self.ctx, self.A, self.grid, self.terms, self.phi, self.rhs, self.x)
        if base:
            _base_guess(self)

    def advance(self, x_solution):
        """phi^{n-1} <- phi^n, phi^n <- solution of the step"""
        from . import copy_to_field
        self.phi_prev.copy_(self.phi)
        copy_to_field(self.ctx, self.grid, x_solution, self.phi)
        self.step_index += 1


class VarDiffusion(_BaseGuess):
    """Transient variable-diffusivity diffusion on an n^3 unit cube (the precision/learned target).

    Per seed: cell diffusivity ln D ~ N(ln 1e-3, 1) (face value: harmonic mean), source
    U[0, 1) per cell (rhs += V s), Dirichlet 0 walls, initial field a sum of 8 Gaussians, and
    dt chosen so that the mean of D dt / dx^2 is `mean_diff_number`. The network sees ln D as a
    fourth input channel ((ln D - ln 1e-3) / 1).
    """

    @staticmethod
    def in_channels_for():
        return 4


    def __init__(self, ctx, seed, n=256, mean_diff_number=1.0, A=None, device="cuda"):
        from . import Csr, build_structure
        torch = _torch()
        rng = np.random.default_rng(seed)
        self.ctx, self.n, self.seed = ctx, n, seed
        self.D, self.src, self.dt = c3_lognormal(n, device=device, seed=seed, mean_diff_number=mean_diff_number)
        self.lnD = torch.log(self.D)
        self.grid = Grid.make([n, n, n], size=[1.0, 1.0, 1.0], ghost=2)
        self.terms = Terms.make(dt=self.dt, diff_cell=self.D, source=self.src, bc_type=[BC_DIRICHLET] * 6,
                                bc_value=[0.0] * 6)
        self.phi = gaussian_sum(n, rng, device=device)
        self.ghost_modes, self.ghost_values = [GHOST_CONSTANT] * 6, [0.0] * 6
        self.phi_prev = self.phi.clone()
        self.rhs = torch.zeros(n ** 3, dtype=torch.float64, device=device)
        self.x = torch.zeros_like(self.rhs)
        self.r0 = torch.zeros_like(self.rhs)
        self.xb = torch.zeros_like(self.rhs)
        self.A = A
        if self.A is None:
            self.A = Csr(ctx)
            build_structure(ctx, self.A, self.grid, self.terms)
        self.rhs_scale = self.dt / (1.0 / n) ** 3
        self.step_index = 0
        self.extras = [(self.lnD, 1.0, math.log(1e-3))]
        self.in_channels = 4

    def assemble(self, base=True):
        from . import fill_values, ghost_fill
        ghost_fill(self.ctx, self.grid, self.phi, self.ghost_modes, self.ghost_values)
        fill_values(self.ctx, self.A, self.grid, self.terms, self.phi, self.rhs, self.x)
        if base:
            _base_guess(self)

    def advance(self, x_solution):
        from . import copy_to_field
        self.phi_prev.copy_(self.phi)
        copy_to_field(self.ctx, self.grid, x_solution, self.phi)
        self.step_index += 1


WORKLOADS = {"advdiff": C5AdvDiff, "vardiff": VarDiffusion}
