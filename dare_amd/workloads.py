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
