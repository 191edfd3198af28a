"""Batched solves (dare_batch_*) against the single-system solver on the same systems.

The batch runs the same BiCGStab2 pass for every system with a per-system mask, so each
system's pass count, stop decision and solution must agree with dare_bicgstab2_solve on that
system alone (reductions are summed in another order: solutions within 1e-10 relative at a
tight tolerance, pass counts within one). Graph capture and launch mode must not change the
results at all; the mixed form must reach the FP64 true residual per system.
"""
import ctypes

import numpy as np
import pytest

import dare_amd as da

pytestmark = pytest.mark.gpu


def make_systems(ctx, rng, dims_list, ncomp, react=True, dt=1e-2):
    import torch
    out = []
    for n in dims_list:
        g = da.Grid.make(list(n), size=[v / 32 for v in n], ghost=1, ncomp=ncomp)
        ncell = int(np.prod(n))
        kw = dict(dt=dt, bc_type=[da.BC_DIRICHLET] * 6, bc_value=[0.0] * 6)
        R = None
        if ncomp == 2:
            kw["diff_comp"] = list(rng.uniform(1e-4, 1e-2, size=2))
            if react:
                k = rng.uniform(0.1, 10, size=(ncell, 2))
                R = torch.from_numpy(np.stack([-k[:, 0], k[:, 1], k[:, 0], -k[:, 1]], 1).ravel()).cuda()
                kw["react"] = R
        else:
            kw["diff"] = float(rng.uniform(1e-4, 1e-2))
        t = da.Terms.make(**kw)
        A = da.Csr(ctx)
        da.build_structure(ctx, A, g, t)
        npad = int(np.prod([v + 2 for v in n])) * ncomp
        phi = torch.from_numpy(rng.uniform(size=npad)).cuda()
        out.append(dict(g=g, t=t, A=A, phi=phi, R=R, rows=ncell * ncomp))
    return out


def d2h(ctx, addr, n):
    a = np.empty(n, np.float64)
    da.check(da.lib().dare_memcpy_d2h(ctx.h, a.ctypes.data, addr, 8 * n), "d2h")
    return a


def fill_batch(ctx, B, systems):
    for s, S in enumerate(systems):
        da.fill_values(ctx, S["A"], S["g"], S["t"], S["phi"], B.rhs_ptr(s), B.x_ptr(s))
    ctx.synchronize()


DIMS = [(5, 7, 9), (16, 3, 20), (12, 12, 12), (3, 4, 70), (33, 9, 17), (8, 40, 6), (20, 20, 33), (4, 4, 4),
        (2, 30, 31), (17, 17, 65)]


@pytest.mark.parametrize("ncomp", [1, 2])
def test_batch_fp64_matches_single_solves(ctx, rng, ncomp):
    systems = make_systems(ctx, rng, DIMS, ncomp)
    B = da.Batch(ctx, [S["A"] for S in systems])
    fill_batch(ctx, B, systems)
    res = B.solve(tol=1e-12, tol_kind=1, max_it=500, graph_passes=3)
    singles = single_solves_rel(ctx, systems, 1e-12, 500)
    for s, (r, (rs, xs, _)) in enumerate(zip(res, singles)):
        assert r["converged"] == 1 and rs["converged"] == 1, s
        assert abs(r["iterations"] - rs["iterations"]) <= 1, (s, r["iterations"], rs["iterations"])
        xb = d2h(ctx, B.x_ptr(s), systems[s]["rows"])
        assert np.max(np.abs(xb - xs)) <= 1e-10 * np.max(np.abs(xs)), s
    B.close()


def single_solves_rel(ctx, systems, tol, max_it):
    import torch
    out = []
    for S in systems:
        b = torch.zeros(S["rows"], dtype=torch.float64, device="cuda")
        x = torch.zeros_like(b)
        da.fill_values(ctx, S["A"], S["g"], S["t"], S["phi"], b, x)
        r = da.bicgstab2(ctx, S["A"], b, x, tol=tol, tol_kind=1, max_it=max_it)
        out.append((r, x.cpu().numpy(), b))
    return out


def test_batch_graph_and_launch_modes_identical(ctx, rng):
    systems = make_systems(ctx, rng, DIMS[:6], 2)
    B = da.Batch(ctx, [S["A"] for S in systems])
    xs = []
    for passes in (0, 1, 4, 7):
        fill_batch(ctx, B, systems)
        res = B.solve(tol=1e-10, max_it=300, graph_passes=passes)
        xs.append((d2h(ctx, B.x_addr, B.rows), [r["iterations"] for r in res]))
    for x, it in xs[1:]:
        assert np.array_equal(x.view(np.int64), xs[0][0].view(np.int64))
        assert it == xs[0][1]
    B.close()


def test_batch_masks_limits_and_zero_rhs(ctx, rng):
    """A system with b = 0 (x0 = 0) stops at once, and a pass limit stops the others, each
    with the single solver's pass count, convergence flag and status (max_it + 1 passes when
    the limit is hit)."""
    import torch
    systems = make_systems(ctx, rng, DIMS[:5], 2)
    B = da.Batch(ctx, [S["A"] for S in systems])
    fill_batch(ctx, B, systems)
    # zero the rhs and x of system 2
    n2 = systems[2]["rows"]
    z = np.zeros(n2)
    da.check(da.lib().dare_memcpy_h2d(ctx.h, B.rhs_ptr(2), z.ctypes.data, 8 * n2), "h2d")
    da.check(da.lib().dare_memcpy_h2d(ctx.h, B.x_ptr(2), z.ctypes.data, 8 * n2), "h2d")
    res = B.solve(tol=1e-14, max_it=3, graph_passes=2)
    for s in range(5):
        S = systems[s]
        b = torch.zeros(S["rows"], dtype=torch.float64, device="cuda")
        x = torch.zeros_like(b)
        if s != 2:
            da.fill_values(ctx, S["A"], S["g"], S["t"], S["phi"], b, x)
        r = da.bicgstab2(ctx, S["A"], b, x, tol=1e-14, tol_kind=1, max_it=3)
        got = (res[s]["converged"], res[s]["iterations"], res[s]["status"])
        assert got == (r["converged"], r["iterations"], r["status"]), (s, got, r)
    assert res[2]["iterations"] <= 1 and res[2]["converged"] == 1
    assert all(res[s]["iterations"] == 4 and not res[s]["converged"] for s in (0, 1, 3, 4))
    B.close()


def test_batch_mixed_reaches_fp64_residual(ctx, rng):
    import torch
    systems = make_systems(ctx, rng, DIMS, 2)
    B = da.Batch(ctx, [S["A"] for S in systems])
    fill_batch(ctx, B, systems)
    res = B.solve(tol=1e-10, tol_kind=1, mixed=True, inner_tol=1e-4, graph_passes=4)
    for s, (S, r) in enumerate(zip(systems, res)):
        assert r["converged"] == 1, (s, r)
        assert r["refinements"] >= 1
        x = torch.from_numpy(d2h(ctx, B.x_ptr(s), S["rows"])).cuda()
        b = torch.from_numpy(d2h(ctx, B.rhs_ptr(s), S["rows"])).cuda()
        y = torch.empty_like(x)
        da.spmv(ctx, S["A"], x, y)
        rel = float(torch.linalg.norm(b - y) / torch.linalg.norm(b))
        assert rel <= 1e-10, (s, rel)
        assert abs(r["residual"] - float(torch.linalg.norm(b - y))) <= 1e-3 * r["residual"] + 1e-300
    B.close()


def test_batch_autotune_and_configs(ctx, rng, tmp_path):
    from dare_amd import autotune
    systems = make_systems(ctx, rng, DIMS, 2)
    B = da.Batch(ctx, [S["A"] for S in systems])
    fill_batch(ctx, B, systems)
    base = B.solve(tol=1e-12, max_it=500, graph_passes=4)
    x0 = d2h(ctx, B.x_addr, B.rows)
    cache = str(tmp_path / "tune.json")
    cfg, us = autotune.tune(B, [S["g"].n[:3] for S in systems], 2, "test", cache_path=cache, reps=2)
    assert us is not None and us > 0 and cfg["grid"] % 8 == 0
    cfg2, us2 = autotune.tune(B, [S["g"].n[:3] for S in systems], 2, "test", cache_path=cache)
    assert cfg2 == cfg and us2 is None and B.config() == cfg
    for c in ({"tile_k": 16, "tile_i": 1, "grid": 8, "order": 0}, {"tile_k": 64, "tile_i": 64, "grid": 4096, "order": 2}):
        B.set_config(**c)
        fill_batch(ctx, B, systems)
        res = B.solve(tol=1e-12, max_it=500, graph_passes=4)
        x = d2h(ctx, B.x_addr, B.rows)
        assert all(r["converged"] for r in res)
        assert [abs(r["iterations"] - q["iterations"]) <= 1 for r, q in zip(res, base)] == [True] * len(res)
        assert np.max(np.abs(x - x0)) <= 1e-10 * np.max(np.abs(x0))
    B.close()


def test_batch_rejects_mixed_components_and_unstructured(ctx, rng):
    s1 = make_systems(ctx, rng, DIMS[:1], 1)
    s2 = make_systems(ctx, rng, DIMS[:1], 2)
    with pytest.raises(da.DareError):
        da.Batch(ctx, [s1[0]["A"], s2[0]["A"]])
    A = da.Csr(ctx)
    da.upload_host_csr(ctx, A, np.array([0, 1]), np.array([0]), np.array([2.0]))
    with pytest.raises(da.DareError):
        da.Batch(ctx, [A])
