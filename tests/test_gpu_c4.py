"""The batching form's own workload against the oracle.

C4-type systems: two components, a 2x2 linearised reaction block [[-k1, k2], [k1, -k2]] per
cell (k1, k2 ~ U[0.1, 10]), D1, D2 ~ U[1e-4, 1e-2], edge lengths log-uniform in [16, 96],
isotropic cells of width 1/64, implicit Euler, Dirichlet faces, two ghost layers. Each batched
solution is compared with the oracle's assembly + BiCGStab2 of the same system
(BiCGStab2.inl:167-582; the cross-component entries of test_TrilinosSolver.cpp:157-164):
FP64 solutions within 1e-10 relative with pass counts within one, mixed solutions at the FP64
true residual computed by the oracle's SpMV. The full 512-system batch must converge in every
system to true relative residual 1e-8.
"""
import math

import numpy as np
import pytest

import dare_amd as da
import oracle as orc

pytestmark = pytest.mark.gpu

H = 1.0 / 64


def c4_systems(rng, nsys):
    dims = np.clip(np.rint(np.exp(rng.uniform(math.log(16), math.log(96), size=(nsys, 3)))), 16, 96).astype(int)
    out = []
    for s in range(nsys):
        n = [int(v) for v in dims[s]]
        ncell = n[0] * n[1] * n[2]
        k1, k2 = rng.uniform(0.1, 10, ncell), rng.uniform(0.1, 10, ncell)
        R = np.stack([-k1, k2, k1, -k2], 1).ravel().copy()
        D = list(rng.uniform(1e-4, 1e-2, size=2))
        bv = list(rng.uniform(-1, 1, size=6))
        phi = rng.uniform(size=(n[0] + 4) * (n[1] + 4) * (n[2] + 4) * 2)
        out.append(dict(n=n, R=R, D=D, bv=bv, phi=phi, rows=2 * ncell))
    return out


def oracle_system(S, dt=1e-2):
    go = orc.grid(S["n"], size=[v * H for v in S["n"]], ghost=2, ncomp=2)
    t = orc.terms(dt=dt, diff_comp=S["D"], react=S["R"], bc_type=[1] * 6, bc_value=S["bv"])
    return go, orc.assemble(go, t, S["phi"])


def device_batch(ctx, systems, dt=1e-2):
    import torch
    keep, csrs = [], []
    for S in systems:
        g = da.Grid.make(S["n"], size=[v * H for v in S["n"]], ghost=2, ncomp=2)
        R = torch.from_numpy(S["R"]).cuda()
        t = da.Terms.make(dt=dt, diff_comp=S["D"], react=R, bc_type=[da.BC_DIRICHLET] * 6, bc_value=S["bv"])
        A = da.Csr(ctx)
        da.build_structure(ctx, A, g, t)
        phi = torch.from_numpy(S["phi"]).cuda()
        keep.append((g, t, A, phi, R))
        csrs.append(A)
    B = da.Batch(ctx, csrs)

    def fill():
        for s, (g, t, A, phi, R) in enumerate(keep):
            da.fill_values(ctx, A, g, t, phi, B.rhs_ptr(s), B.x_ptr(s))
        ctx.synchronize()

    return B, keep, fill


def d2h(ctx, addr, n):
    a = np.empty(n, np.float64)
    da.check(da.lib().dare_memcpy_d2h(ctx.h, a.ctypes.data, addr, 8 * n), "d2h")
    return a


@pytest.fixture(scope="module")
def c4_eight(ctx):
    orc.set_threads(min(16, __import__("os").cpu_count() or 1))
    rng = np.random.default_rng(2026)
    systems = c4_systems(rng, 8)
    refs = []
    for S in systems:
        go, (rp, col, val, rhs, x0) = oracle_system(S)
        tol = 1e-12 * np.linalg.norm(rhs)
        xo, ok, it, res, _ = orc.bicgstab2(rp, col, val, rhs, x0, tol, 500)
        assert ok
        refs.append(dict(rp=rp, col=col, val=val, rhs=rhs, x0=x0, x=xo, it=it))
    print("C4 systems:", [S["n"] for S in systems], flush=True)
    return systems, refs


def test_c4_batch_fp64_matches_oracle(ctx, c4_eight):
    systems, refs = c4_eight
    B, keep, fill = device_batch(ctx, systems)
    fill()
    for s, ref in enumerate(refs):     # assembly into the batch's vectors: rhs and x0 bit-exact
        assert np.array_equal(d2h(ctx, B.rhs_ptr(s), len(ref["rhs"])).view(np.int64), ref["rhs"].view(np.int64))
        assert np.array_equal(d2h(ctx, B.x_ptr(s), len(ref["x0"])).view(np.int64), ref["x0"].view(np.int64))
    res = B.solve(tol=1e-12, tol_kind=1, max_it=500, mixed=False, graph_passes=3)
    for s, (r, ref) in enumerate(zip(res, refs)):
        assert r["converged"] == 1, s
        assert abs(r["iterations"] - ref["it"]) <= 1, (s, r["iterations"], ref["it"])
        x = d2h(ctx, B.x_ptr(s), len(ref["x"]))
        assert np.max(np.abs(x - ref["x"])) <= 1e-10 * np.max(np.abs(ref["x"])), s
    B.close()
    for k in keep:
        k[2].close()


def test_c4_batch_mixed_reaches_oracle_true_residual(ctx, c4_eight):
    systems, refs = c4_eight
    B, keep, fill = device_batch(ctx, systems)
    fill()
    res = B.solve(tol=1e-10, tol_kind=1, max_it=1000, mixed=True, inner_tol=1e-4, graph_passes=3)
    for s, (r, ref) in enumerate(zip(res, refs)):
        assert r["converged"] == 1, s
        x = d2h(ctx, B.x_ptr(s), len(ref["x"]))
        true_rel = np.linalg.norm(ref["rhs"] - orc.spmv(ref["rp"], ref["col"], ref["val"], x)) / np.linalg.norm(ref["rhs"])
        assert true_rel <= 1e-10, (s, true_rel)
        assert np.max(np.abs(x - ref["x"])) <= 1e-8 * np.max(np.abs(ref["x"])), s
    B.close()
    for k in keep:
        k[2].close()


def test_c4_full_batch_converges_every_system(ctx):
    """The 512-system C4 workload (dare_amd.workloads.C4Batch): mixed batched solve to 1e-8.
    Every system's true relative residual comes from the contract CSR SpMV (dare_spmv, bit-exact
    against the oracle's SpMV in test_gpu_parity) on the device; on four sampled systems the
    oracle's own SpMV recomputes it on the host."""
    import torch
    from dare_amd.workloads import C4Batch
    W = C4Batch(ctx, nsys=512)
    W.assemble()
    B = W.batch
    b_all = torch.empty(B.rows, dtype=torch.float64, device="cuda")
    da.check(da.lib().dare_memcpy_d2d(ctx.h, b_all.data_ptr(), B.b_addr, 8 * B.rows), "copy b")
    res = W.batch.solve(tol=1e-8, tol_kind=1, max_it=1000, mixed=True, inner_tol=1e-4, graph_passes=3)
    assert sum(r["converged"] for r in res) == 512
    offs = B.offsets
    worst = 0.0
    sample = set(np.random.default_rng(1).choice(512, 4, replace=False).tolist())
    for s in range(512):
        rows = int(offs[s + 1] - offs[s])
        x = torch.empty(rows, dtype=torch.float64, device="cuda")
        da.check(da.lib().dare_memcpy_d2d(ctx.h, x.data_ptr(), B.x_ptr(s), 8 * rows), "copy x")
        y = torch.empty_like(x)
        da.spmv(ctx, W.csr[s], x, y)
        b = b_all[int(offs[s]):int(offs[s + 1])]
        rel = float(torch.linalg.norm(b - y) / torch.linalg.norm(b))
        worst = max(worst, rel)
        if s in sample:
            rp, col, val = W.csr[s].download()
            xh, bh = x.cpu().numpy(), b.cpu().numpy()
            rel_o = np.linalg.norm(bh - orc.spmv(rp, col, val, xh)) / np.linalg.norm(bh)
            assert abs(rel_o - rel) <= 1e-6 * rel + 1e-300, (s, rel, rel_o)
    print(f"C4 512 systems: worst true relative residual {worst:.3e}", flush=True)
    assert worst <= 1e-8
    W.close()
