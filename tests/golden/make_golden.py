"""Writes the golden fixtures of the hot path (tests/golden/golden_8.npz, golden_16.npz).

    python tests/golden/make_golden.py

Inputs come from numpy's seeded PCG64 generator; outputs from the CPU oracle (oracle/), whose
functions are pinned in tests/test_oracle.py and tests/test_oracle_pins.py. The fixtures let
the device path be checked without a live oracle build: tests/test_golden.py regenerates them
with the oracle (bit for bit) and tests/test_gpu_golden.py runs the HIP path against them.

Contents per grid (n^3 internal cells, two ghost layers, unit cube):
  phi, vel_x/y/z              operator-comparison field U[0,1) with flat patches, face velocities
  grad_x/y/z, div             face gradient of phi, value divergence of the gradient
  tvd_<lim>_x/y/z             TVD face fluxes phi_f u_f per limiter
  step_vanleer                explicit van Leer TVD step, dt = 1e-3
  ghost                       ghost fill, modes (Dirichlet, Neumann, constant) per face
  ucell_x/y/z, uface_x/y/z    cell-centred velocity and its face interpolation
  <op>_rp/col/val/rhs/x0      CSR assembly of an operator (see OPERATORS)
  <op>_hist, <op>_x, <op>_it  BiCGStab2 residual norms after passes 1-3, solution, passes
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import oracle as orc  # noqa: E402

from spec import GHOST_MODES, GHOST_VALUES, LIMITERS, operators  # noqa: E402


def make(n, seed):
    rng = np.random.default_rng(seed)
    g = orc.grid([n, n, n], ghost=2)
    npad = (n + 4) ** 3
    out = {}
    phi = rng.uniform(size=npad)
    phi[rng.uniform(size=npad) < 0.3] = 1.0
    out["phi"] = phi
    vel = []
    for d, ax in enumerate("xyz"):
        u = rng.uniform(-1, 1, size=orc.lib().orc_num_faces(ctypes.byref(g), d))
        u[rng.uniform(size=u.size) < 0.05] = 0.0
        vel.append(u)
        out[f"vel_{ax}"] = u
    grad = orc.grad_faces(g, phi)
    for d, ax in enumerate("xyz"):
        out[f"grad_{ax}"] = grad[d]
    out["div"] = orc.div_faces(g, grad)
    lims = LIMITERS if n == 8 else {"vanleer": 4}
    for name, lid in lims.items():
        f = orc.tvd_face_flux(g, lid, phi, vel)
        for d, ax in enumerate("xyz"):
            out[f"tvd_{name}_{ax}"] = f[d]
    out["step_vanleer"] = orc.tvd_explicit_step(g, 4, phi, vel, 1e-3)
    out["ghost"] = orc.ghost_fill(g, phi, GHOST_MODES, GHOST_VALUES)
    if n == 8:
        uc = [rng.uniform(-1, 1, size=npad) for _ in range(3)]
        uf = orc.interpolate_to_faces(g, uc)
        for d, ax in enumerate("xyz"):
            out[f"ucell_{ax}"], out[f"uface_{ax}"] = uc[d], uf[d]
    for name, ncomp, kw, extra in operators(n):
        go = orc.grid([n, n, n], ghost=2, ncomp=ncomp)
        arrays = {}
        if "diff_cell" in extra:
            arrays["diff_cell"] = out[f"{name}_diffcell"] = np.exp(rng.normal(np.log(1e-3), 1.0, size=npad))
        if "source" in extra:
            arrays["source"] = out[f"{name}_source"] = rng.uniform(size=npad)
        if "vel" in extra:
            arrays["vel"] = vel
        if "react" in extra:
            k1, k2 = rng.uniform(0.1, 10, n ** 3), rng.uniform(0.1, 10, n ** 3)
            arrays["react"] = out[f"{name}_react"] = np.stack([-k1, k2, k1, -k2], 1).ravel().copy()
        field = phi if ncomp == 1 else out.setdefault(f"{name}_phi", rng.uniform(size=npad * ncomp))
        t = orc.terms(**kw, **arrays)
        rp, col, val, rhs, x0 = orc.assemble(go, t, field)
        tol = 1e-12 * np.linalg.norm(rhs)
        x, ok, it, res, hist = orc.bicgstab2(rp, col, val, rhs, x0, tol, 500, 3)
        assert ok
        out.update({f"{name}_rp": rp, f"{name}_col": col, f"{name}_val": val, f"{name}_rhs": rhs,
                    f"{name}_x0": x0, f"{name}_x": x, f"{name}_hist": hist, f"{name}_it": np.array([it]),
                    f"{name}_tol": np.array([tol])})
    return out


def generate(n):
    return make(n, seed=1000 + n)


def main():
    for n in (8, 16):
        data = generate(n)
        path = os.path.join(HERE, f"golden_{n}.npz")
        np.savez_compressed(path, **data)
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
