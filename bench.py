"""Benchmark of the FV hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 256^3 convection-diffusion.
One step = ghost fill + explicit TVD advection (van Leer, CFL 0.4, solid-body rotation) +
assembly of the implicit diffusion operator (D = 1e-3, Dirichlet 0) into the device CSR +
BiCGStab2 to relative residual 1e-8 + copy-back. Inputs are resident in HBM before timing.

Prints ONE JSON line. `value` is the implicit solve time (assembly + Krylov solve to 1e-8)
per step of the fastest form; the plain FP64 form and the mixed-precision form are both
reported. `roofline` is measured live on the dominant kernel (fused SpMV+dot over the FP64
diagonal view of the CSR matrix) with
HIP events on the stream the kernel runs on. `cpu_baseline` is the oracle (a C++ restatement
of the reference algorithm) run on the host cores for one full 256^3 implicit step.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--n 256] [--no-cpu]
"""
import argparse
import json
import subprocess
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TOL = 1e-8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3 (no CPU leg)")
    ap.add_argument("--no-extra", action="store_true", help="skip the target / batch / learned legs")
    ap.add_argument("--target-steps", type=int, default=4)
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--only", choices=["target", "precision", "batch", "learned"], default=None,
                    help="run one extra leg alone and print its JSON (profiling)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this implementation has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import dare_amd as da
    from dare_amd.workloads import C2Rotation

    torch.cuda.set_device(0)
    ctx = da.Context(0)
    if args.only:
        leg = {"target": lambda: target_leg(ctx, args.n, args.target_steps), "precision": lambda: precision_leg(ctx),
               "batch": lambda: batch_leg(ctx, args.batch_systems), "learned": lambda: learned_leg(ctx)}[args.only]
        print(json.dumps({args.only: leg()}), flush=True)
        ctx.close()
        return
    W = C2Rotation(args.n)
    g = W.grid
    log(f"[bench] grid {args.n}^3, dt {W.dt:.3e}, building structure")
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, W.terms)
    ctx.synchronize()
    nrows, nnz = A.nrows, A.nnz

    def step(form, phi, stats):
        da.ghost_fill(ctx, g, phi, W.ghost_modes, W.ghost_values)
        da.tvd_explicit_step(ctx, g, da.VANLEER, phi, W.vel, W.dt, W.phi_work)
        ctx.synchronize()
        t0 = time.perf_counter()
        da.fill_values(ctx, A, g, W.terms, W.phi_work, W.rhs, W.x)
        if form == "fp64":
            r = da.bicgstab2(ctx, A, W.rhs, W.x, tol=TOL, tol_kind=1, max_it=1000)
        else:
            r = da.refine_mixed(ctx, A, W.rhs, W.x, tol=TOL, tol_kind=1, inner_tol=1e-4,
                                max_refine=30, inner_max_it=1000, max_it=1000)
        ctx.synchronize()
        stats["solve_s"].append(time.perf_counter() - t0)
        stats["iters"].append(r["iterations"])
        stats["conv"].append(r["converged"])
        da.copy_to_field(ctx, g, W.x, phi)
        return r

    def run(form, steps, warmup):
        phi = W.phi.clone()
        stats = {"solve_s": [], "iters": [], "conv": []}
        for _ in range(warmup):
            step(form, phi, stats)
        stats = {"solve_s": [], "iters": [], "conv": []}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(form, phi, stats)
        torch.cuda.synchronize()
        total = time.perf_counter() - t0
        return total, stats, phi

    results = {}
    for form in ("fp64", "mixed"):
        log(f"[bench] form {form}: warmup {args.warmup}, steps {args.steps}")
        total, st, phi = run(form, args.steps, args.warmup)
        # true FP64 relative residual of the last solve
        y = torch.empty_like(W.x)
        da.spmv(ctx, A, W.x, y)
        torch.cuda.synchronize()
        true_rel = float(torch.linalg.norm(W.rhs - y) / torch.linalg.norm(W.rhs))
        results[form] = dict(ms_per_step=1e3 * total / args.steps,
                             solve_ms=1e3 * sum(st["solve_s"]) / len(st["solve_s"]),
                             iterations=st["iters"], converged=all(st["conv"]),
                             true_rel_residual=true_rel)
        log(f"[bench] {form}: {results[form]}")

    # live roofline of the dominant kernel: the solver's fused SpMV+dot over the diagonal view,
    # HIP events on the ctx stream (the context follows torch's current stream). The C2
    # operator is exactly symmetric, so the solver runs the four-slot traversal: 4 FP64 slots +
    # x + y + the dot operand = 56 B/row; the seven-slot traversal of a general operator (80
    # B/row) is measured on the same matrix with the symmetric traversal switched off. Measured
    # as the solver runs it: before each timed launch the input vector is rewritten with
    # non-temporal stores (the library's copy probe), as the vector update before every SpMV of
    # a pass leaves it; back-to-back launches on an untouched input (isolated_avg_launch_us)
    # run faster.
    x = W.x
    y = torch.empty_like(x)
    w = W.rhs
    xsave = x.clone()
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def spmv_times():
        for _ in range(3):
            da.spmv_dot(ctx, A, x, y, w, readback=False)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            da.spmv_dot(ctx, A, x, y, w, readback=False)
        e1.record()
        torch.cuda.synchronize()
        iso = e0.elapsed_time(e1) / 1e3 / reps
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in evs:
            da.stream_probe(ctx, xsave, x, ratio=1)
            a.record()
            da.spmv_dot(ctx, A, x, y, w, readback=False)
            b.record()
        torch.cuda.synchronize()
        return iso, sum(a.elapsed_time(b) for a, b in evs) / 1e3 / reps

    sym = da.csr_symmetric(ctx, A)
    spmv_iso_s, spmv_s = spmv_times()
    da.set_symmetric_traversal(A, False)
    gen_iso_s, gen_s = spmv_times()
    da.set_symmetric_traversal(A, True)
    del xsave
    spmv_bpr = (4 if sym else 7) * 8 + 8 + 8 + 8
    spmv_bytes = spmv_bpr * nrows
    gen_bytes = 80 * nrows
    achieved = spmv_bytes / spmv_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "spmv_dot_dia_sym_f64_256_traffic.json" if sym
                         else "spmv_dot_dia_f64_256_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except Exception:
            traffic = None
    # attainable HBM bandwidth on this device, same timing, from the library's own streaming
    # probe (dare_stream_probe): a 1 GiB copy, a read-heavy mix (8 reads per write, 1 GiB read,
    # eight separate arrays like the diagonal view's slots) near the SpMV's 9:1, and 1 GiB read
    # only (the ceiling of a read-dominated kernel); torch's copy and fill for comparison
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    src.fill_(1.0)
    dst = torch.empty_like(src)
    dst8 = torch.empty(src.numel() // 8, dtype=torch.float64, device="cuda")

    def probe(fn, nbytes):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return nbytes / (e0.elapsed_time(e1) / 1e3 / reps) / 1e9

    stream_copy_gbs = probe(lambda: da.stream_probe(ctx, src, dst, ratio=1), 2 * src.numel() * 8)
    stream_read8_gbs = probe(lambda: da.stream_probe(ctx, src, dst8, ratio=8, layout=2),
                             (src.numel() + dst8.numel()) * 8)
    ro_grid = 8 * torch.cuda.get_device_properties(0).multi_processor_count
    stream_read_gbs = probe(lambda: da.stream_probe(ctx, src, dst8, ratio=8, grid=ro_grid, layout=3),
                            src.numel() * 8 + ro_grid * 256 * 16)
    copy_gbs = probe(lambda: dst.copy_(src), 2 * src.numel() * 8)
    write_gbs = probe(lambda: dst.fill_(1.0), dst.numel() * 8)
    del src, dst, dst8

    # the contract-format SpMV (dare_spmv on a CSR without a diagonal view: k_apply_csr), on the
    # same 256^3 operator uploaded as a plain CSR: 12 nnz + 4 (n + 1) + 8 n (x) + 8 n (y) bytes
    rp_h, col_h, val_h = A.download()
    C = da.Csr(ctx)
    da.upload_host_csr(ctx, C, rp_h, col_h, val_h)
    del rp_h, col_h, val_h
    yc = torch.empty_like(x)

    # the stencil kernels, same method: TVD step (phi + 3 face velocities in, phi out), face
    # gradient (phi in, 3 face arrays out), value divergence (3 face arrays in, cells out)
    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    # operator-comparison field: phi ~ U[0,1) on every cell, ghosts included (seeded); the
    # kernels have no data-dependent paths, and this field leaves nothing flat to exploit
    ncell = args.n ** 3
    nface = 3 * ncell + 3 * args.n ** 2
    gen = torch.Generator(device="cuda")
    gen.manual_seed(2024)
    phi_u = torch.rand(W.phi.numel(), generator=gen, dtype=torch.float64, device="cuda")
    csr_s = timed(lambda: da.spmv(ctx, C, x, yc))
    C.close()
    del yc
    tvd_s = timed(lambda: da.tvd_explicit_step(ctx, g, da.VANLEER, phi_u, W.vel, W.dt, W.phi_work))
    tvd_bytes = 8 * (2 * ncell + nface)
    faces = [torch.empty(v.numel(), dtype=torch.float64, device="cuda") for v in W.vel]
    grad_s = timed(lambda: da.grad_faces(ctx, g, phi_u, faces))
    div_s = timed(lambda: da.div_faces(ctx, g, faces, W.phi_work))
    del faces, phi_u
    stencils = {}
    csr_bytes = 12 * nnz + 4 * (nrows + 1) + 16 * nrows
    for name, kern, sec, nbytes in (("tvd_step", "k_tvd_step3", tvd_s, tvd_bytes),
                                    ("gradient", "k_grad3_pair", grad_s, 8 * (ncell + nface)),
                                    ("divergence", "k_div3", div_s, 8 * (nface + ncell)),
                                    ("csr_spmv", "k_apply_csr<double,double,EpiStore>", csr_s, csr_bytes)):
        gbs = nbytes / sec / 1e9
        stencils[name] = {"kernel": kern, "bytes_per_cell": round(nbytes / ncell, 3), "achieved": round(gbs, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                          "avg_launch_us": round(sec * 1e6, 2), "cell_updates_per_s": round(ncell / sec, 1)}

    best = min(results, key=lambda f: results[f]["solve_ms"] if results[f]["converged"] else 1e30)

    cpu = None
    if not args.no_cpu and not args.profile:
        cpu = cpu_baseline(args, W, ctx)
    extra = {}
    if not args.no_extra and not args.profile:
        A.close()
        del W
        torch.cuda.empty_cache()
        extra["target"] = target_leg(ctx, args.n, args.target_steps)
        torch.cuda.empty_cache()
        extra["precision"] = precision_leg(ctx)
        torch.cuda.empty_cache()
        extra["batch"] = batch_leg(ctx, args.batch_systems)
        torch.cuda.empty_cache()
        extra["learned"] = learned_leg(ctx)

    line = {
        "metric": "256^3 solve time to 1e-8 resid: mixed/learned vs plain FP64 GPU; HBM GB/s",
        "value": round(results[best]["solve_ms"], 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(results[best]["ms_per_step"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else "f64 (f32 inner)",
        "data": "synthetic",
        "config": {
            "workload": f"{args.n}^3 convection-diffusion step: TVD van Leer advection (CFL 0.4, "
                        f"solid-body rotation) + implicit diffusion D=1e-3 Dirichlet 0, "
                        f"BiCGStab2 to rel residual 1e-8",
            "grid": [args.n] * 3, "rows": nrows, "nnz": nnz, "form": best,
        },
        "forms": {f: {k: (round(v, 4) if isinstance(v, float) and k.endswith("ms") else v) for k, v in r.items()}
                  for f, r in results.items()},
        "speedup_vs_fp64": round(results["fp64"]["solve_ms"] / results[best]["solve_ms"], 3),
        "roofline": {
            "kernel": ("k_apply_dia<1,double,double,EpiDot,2,true> (SpMV over the symmetric diagonal view, "
                       "four slots, + fused dot)" if sym else
                       "k_apply_dia<1,double,double,EpiDot,2> (SpMV over the diagonal view + fused dot)"),
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes": spmv_bytes, "bytes_per_row": spmv_bpr, "avg_launch_us": round(spmv_s * 1e6, 2),
            "isolated_avg_launch_us": round(spmv_iso_s * 1e6, 2),
            "isolated_frac": round(spmv_bytes / spmv_iso_s / 1e9 / HBM_PEAK_GBS, 4),
            "general_operator": {
                "kernel": "k_apply_dia<1,double,double,EpiDot,2> (seven slots; symmetric traversal off)",
                "bytes_per_row": 80, "avg_launch_us": round(gen_s * 1e6, 2),
                "achieved": round(gen_bytes / gen_s / 1e9, 1), "frac": round(gen_bytes / gen_s / 1e9 / HBM_PEAK_GBS, 4),
                "isolated_avg_launch_us": round(gen_iso_s * 1e6, 2),
                "isolated_frac": round(gen_bytes / gen_iso_s / 1e9 / HBM_PEAK_GBS, 4)},
            "attainable": {"probe": "dare_stream_probe (own kernel, 1 GiB read)",
                           "copy_GBs": round(stream_copy_gbs, 1), "read8_write1_GBs": round(stream_read8_gbs, 1),
                           "read_only_GBs": round(stream_read_gbs, 1),
                           "torch_copy_GBs": round(copy_gbs, 1), "torch_fill_GBs": round(write_gbs, 1)},
            "frac_of_attainable": round(achieved / max(stream_read8_gbs, stream_copy_gbs, stream_read_gbs), 4),
        },
        "stencil_rooflines": stencils,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)
    A.close()
    ctx.close()


WEIGHTS = os.path.join(ROOT, "dare_amd", "weights")
WEIGHTS_OVERRIDE = None        # a weights file for the target leg's network (experiments)


def _true_rel(ctx, A, b, x):
    import torch
    import dare_amd as da
    y = torch.empty_like(x)
    da.spmv(ctx, A, x, y)
    return float(torch.linalg.norm(b - y) / torch.linalg.norm(b))


def target_leg(ctx, n, steps, inner_tol=1e-4):
    """256^3 variable-diffusivity implicit steps (ln D ~ N(ln 1e-3, 1), harmonic face values,
    source U[0,1), Dirichlet 0, mean D dt/dx^2 = 1) on a held-out seed. The plain FP64
    BiCGStab2 from phi^n (the reference's method and guess) against mixed precision, Jacobi
    right preconditioning and the extrapolated / learned initial guesses; every form to FP64
    true relative residual 1e-8 (checked). Step 0 (no phi^{n-1}) is a warm-up."""
    import torch
    import dare_amd as da
    from dare_amd.learned import WarmStartNet
    from dare_amd.workloads import VarDiffusion
    log(f"[bench] target: {n}^3 variable-diffusivity steps")
    W = VarDiffusion(ctx, seed=5000, n=n)
    net = None
    # the network fine-tuned at this resolution when there is one (crops of 256^3 trajectories),
    # else the 128^3 one
    wpath = os.path.join(WEIGHTS, f"ws_vardiff_{n}.npz")
    if not os.path.exists(wpath):
        wpath = os.path.join(WEIGHTS, "ws_vardiff_128.npz")
    wpath = WEIGHTS_OVERRIDE or wpath
    if os.path.exists(wpath):
        net = WarmStartNet.load(ctx, wpath, [n] * 3)
    # (name, mixed, precond, guess)
    forms = [("fp64", False, 0, "previous"), ("fp64+jacobi", False, 1, "previous"),
             ("fp64+mg", False, 2, "previous"),
             ("mixed", True, 0, "previous"), ("mixed+jacobi", True, 1, "previous"),
             ("mixed+mg", True, 2, "previous"),
             ("mixed+jacobi+extrapolated", True, 1, "extrapolated"), ("mixed+mg+extrapolated", True, 2, "extrapolated")]
    if net:
        forms += [("mixed+learned", True, 0, "learned"), ("mixed+jacobi+learned", True, 1, "learned"),
                  ("mixed+mg+learned", True, 2, "learned")]
    st = {f[0]: {"ms": [], "passes": [], "refine": [], "rel": [], "guess_ms": [], "rel0": []} for f in forms}
    xg = torch.empty_like(W.x)
    for k in range(steps + 1):
        W.assemble(base=False)
        ctx.synchronize()
        x_prev = W.x.clone()
        keep = None
        for name, mixed, pre, guess in forms:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            # a guess is timed with what it needs: the extrapolation, and for the network also
            # the residual of the extrapolation that it reads
            if guess == "learned":
                W.base_guess()
                net.predict(W.grid, W.phi, W.phi_prev, W.xb, W.r0, W.rhs_scale, xg, extras=W.extras)
            elif guess == "extrapolated":
                W.base_guess(with_residual=False, out=xg)       # formed in the solver's vector
            else:
                xg.copy_(x_prev)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            rel0 = _true_rel(ctx, W.A, W.rhs, xg)   # outside the timed intervals
            torch.cuda.synchronize()
            t1b = time.perf_counter()
            if mixed:
                r = da.refine_mixed(ctx, W.A, W.rhs, xg, tol=TOL, tol_kind=1, inner_tol=inner_tol, max_refine=30,
                                    inner_max_it=2000, max_it=4000, precond=pre)
            else:
                r = da.bicgstab2(ctx, W.A, W.rhs, xg, tol=TOL, tol_kind=1, max_it=2000, precond=pre)
            ctx.synchronize()
            t2 = time.perf_counter()
            if k >= 1:
                st[name]["ms"].append(1e3 * ((t1 - t0) + (t2 - t1b)))
                st[name]["rel0"].append(rel0)
                st[name]["guess_ms"].append(1e3 * (t1 - t0))
                st[name]["passes"].append(r["iterations"])
                st[name]["refine"].append(r["refinements"])
                st[name]["rel"].append(_true_rel(ctx, W.A, W.rhs, xg))
            if name == "fp64":
                keep = xg.clone()
        W.advance(keep)
    out = {"workload": f"{n}^3 variable-diffusivity implicit step (held-out seed 5000), {steps} steps after "
                       f"1 warm-up; guess: previous = phi^n", "dt": W.dt}
    for name, *_ in forms:
        out[name] = {"ms_mean": round(float(np.mean(st[name]["ms"])), 3),
                     "guess_ms_mean": round(float(np.mean(st[name]["guess_ms"])), 3),
                     "passes_mean": float(np.mean(st[name]["passes"])),
                     "refinements_mean": float(np.mean(st[name]["refine"])),
                     "initial_rel_residual_geomean": float(np.exp(np.mean(np.log(st[name]["rel0"])))),
                     "true_rel_residual_max": float(np.max(st[name]["rel"]))}
    base = out["fp64"]["ms_mean"]
    for name, *_ in forms:
        out[name]["speedup_vs_fp64"] = round(base / out[name]["ms_mean"], 3)
    best = max((f[0] for f in forms), key=lambda f: out[f]["speedup_vs_fp64"])
    out["best_form"] = best
    out["best_speedup_vs_fp64"] = out[best]["speedup_vs_fp64"]
    if net:
        net.close()
    W.A.close()
    return out


def precision_leg(ctx, n=384):
    """Precision form at 384^3: one implicit variable-coefficient diffusion step (ln D ~
    N(ln 1e-3, 1), harmonic face values, source U[0,1), Dirichlet 0, mean D dt/dx^2 = 1, from
    phi = 0). FP32 inner BiCGStab2 (inner tolerance 1e-4) inside FP64 refinement to true
    relative residual 1e-10, against the plain FP64 solve to the same tolerance; an FP64 solve
    to 1e-13 is the accuracy reference for both."""
    import torch
    import dare_amd as da
    from dare_amd.workloads import c3_lognormal
    log(f"[bench] precision: {n}^3 variable-coefficient diffusion")
    D, src, dt = c3_lognormal(n, seed=3, mean_diff_number=1.0)
    g = da.Grid.make([n] * 3, size=[1.0, 1.0, 1.0], ghost=2)
    t = da.Terms.make(dt=dt, diff_cell=D, source=src, bc_type=[da.BC_DIRICHLET] * 6, bc_value=[0.0] * 6)
    A = da.Csr(ctx)
    da.build_structure(ctx, A, g, t)
    phi = torch.zeros((n + 4) ** 3, dtype=torch.float64, device="cuda")
    rhs = torch.zeros(n ** 3, dtype=torch.float64, device="cuda")
    x0 = torch.zeros_like(rhs)
    da.fill_values(ctx, A, g, t, phi, rhs, x0)
    ctx.synchronize()
    xr = x0.clone()
    da.bicgstab2(ctx, A, rhs, xr, tol=1e-13, tol_kind=1, max_it=5000, precond=1)
    out = {"workload": f"{n}^3 variable-coefficient diffusion step, to true rel residual 1e-10"}
    for name, mixed, pre in (("fp64", False, 0), ("mixed", True, 0), ("mixed+jacobi", True, 1), ("fp64+mg", False, 2),
                             ("mixed+mg", True, 2)):
        ts = []
        for rep in range(2):
            x = x0.clone()
            ctx.synchronize()
            t0 = time.perf_counter()
            if mixed:
                r = da.refine_mixed(ctx, A, rhs, x, tol=1e-10, tol_kind=1, inner_tol=1e-4, max_refine=30,
                                    inner_max_it=5000, max_it=10000, precond=pre)
            else:
                r = da.bicgstab2(ctx, A, rhs, x, tol=1e-10, tol_kind=1, max_it=5000, precond=pre)
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        err = float(torch.linalg.norm(x - xr) / torch.linalg.norm(xr))
        out[name] = {"ms": round(1e3 * min(ts), 2), "passes": r["iterations"], "refinements": r["refinements"],
                     "converged": r["converged"], "true_rel_residual": _true_rel(ctx, A, rhs, x),
                     "rel_error_vs_fp64_reference": err}
    for name in ("mixed", "mixed+jacobi", "fp64+mg", "mixed+mg"):
        out[name]["speedup_vs_fp64"] = round(out["fp64"]["ms"] / out[name]["ms"], 3)
    A.close()
    return out


def batch_leg(ctx, nsys):
    """512 independent 2-component reaction-diffusion systems (edges log-uniform in [16, 96]):
    one batched solve (persistent launches, per-system masks, graph-captured passes, autotuned
    configuration) against the same systems solved one after the other by the plain solver."""
    import torch
    import dare_amd as da
    from dare_amd import autotune
    from dare_amd.workloads import C4Batch
    log(f"[bench] batch: {nsys} systems")
    t0 = time.perf_counter()
    Wb = C4Batch(ctx, nsys=nsys)
    Wb.assemble()
    B = Wb.batch
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    cfg, us = autotune.tune(B, Wb.dims.tolist(), 2, torch.cuda.get_device_name(0), reps=3)
    tune_s = time.perf_counter() - t0
    x0 = torch.empty(B.rows, dtype=torch.float64, device="cuda")
    b = torch.empty_like(x0)
    da.check(da.lib().dare_memcpy_d2d(ctx.h, x0.data_ptr(), B.x_addr, 8 * B.rows))
    da.check(da.lib().dare_memcpy_d2d(ctx.h, b.data_ptr(), B.b_addr, 8 * B.rows))
    ctx.synchronize()
    res_t = {}
    results = {}
    for mixed in (False, True):
        for rep in range(2):          # first call captures the graphs
            da.check(da.lib().dare_memcpy_d2d(ctx.h, B.x_addr, x0.data_ptr(), 8 * B.rows))
            ctx.synchronize()
            t0 = time.perf_counter()
            res = B.solve(tol=TOL, tol_kind=1, max_it=1000, mixed=mixed, graph_passes=4)
            ctx.synchronize()
            el = time.perf_counter() - t0
        key = "batched_mixed" if mixed else "batched_fp64"
        res_t[key] = el
        results[key] = res
    xb = torch.empty_like(x0)
    da.check(da.lib().dare_memcpy_d2d(ctx.h, xb.data_ptr(), B.x_addr, 8 * B.rows))
    xs = x0.clone()
    ctx.synchronize()
    t0 = time.perf_counter()
    seq = []
    for s in range(nsys):
        o0, o1 = int(B.offsets[s]), int(B.offsets[s + 1])
        seq.append(da.bicgstab2(ctx, Wb.csr[s], b[o0:o1], xs[o0:o1], tol=TOL, tol_kind=1, max_it=1000))
    ctx.synchronize()
    seq_s = time.perf_counter() - t0
    worst = 0.0
    for s in range(nsys):
        o0, o1 = int(B.offsets[s]), int(B.offsets[s + 1])
        worst = max(worst, _true_rel(ctx, Wb.csr[s], b[o0:o1], xb[o0:o1]))
    out = {"workload": f"{nsys} reaction-diffusion systems, N=2, edges log-uniform [16,96], "
                       f"{Wb.cells} cells, {B.rows} rows, tol 1e-8 per system",
           "sequential_plain_ms": round(1e3 * seq_s, 2),
           "batched_fp64_ms": round(1e3 * res_t["batched_fp64"], 2),
           "batched_mixed_ms": round(1e3 * res_t["batched_mixed"], 2),
           "speedup_mixed_vs_sequential": round(seq_s / res_t["batched_mixed"], 2),
           "speedup_fp64_vs_sequential": round(seq_s / res_t["batched_fp64"], 2),
           "converged": {k: int(sum(r["converged"] for r in v)) for k, v in results.items()},
           "sequential_passes_mean": float(np.mean([r["iterations"] for r in seq])),
           "batched_mixed_worst_true_rel_residual": worst,
           "autotuned_config": cfg, "autotune_s": round(tune_s, 2), "build_s": round(build_s, 2)}
    Wb.close()
    return out


def learned_leg(ctx, seeds=2, steps=30):
    """Held-out advection-diffusion trajectories at 128^3: passes and time per step from the
    zero, previous-step, extrapolated and learned initial guesses (a short sample of the full
    evaluation the training report holds)."""
    from dare_amd.learned import WarmStartNet
    from dare_amd.warmstart import evaluate
    wpath = os.path.join(WEIGHTS, "ws_advdiff_128.npz")
    if not os.path.exists(wpath):
        return None
    log("[bench] learned: held-out 128^3 advection-diffusion steps")
    net = WarmStartNet.load(ctx, wpath, [128] * 3)
    ev = evaluate(ctx, net, n=128, eval_seeds=range(5000, 5000 + seeds), steps=steps, workload="advdiff")
    net.close()
    return {"workload": f"128^3 advection-diffusion, {seeds} held-out seeds x {steps} steps", **ev}


def _host_info():
    """lscpu's description of the host (model name, sockets, cores, threads), for the record."""
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keep = {"Model name": "model", "Vendor ID": "vendor", "Socket(s)": "sockets",
                "Core(s) per socket": "cores_per_socket", "Thread(s) per core": "threads_per_core",
                "CPU(s)": "cpus", "NUMA node(s)": "numa_nodes", "CPU max MHz": "max_mhz"}
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() in keep:
                info[keep[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(args, W, ctx):
    """The oracle (C++ restatement of the reference algorithm) on the host: one full implicit
    step of the same workload (assembly + BiCGStab2 to 1e-8 relative). One warm-up step, then
    the median of five timed steps on all host threads (up to 16, the box's share), and the
    median of three on one thread."""
    import numpy as np
    import torch
    import oracle as orc
    import dare_amd as da
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    log(f"[bench] cpu baseline: oracle, {threads} threads, {args.n}^3 implicit step, 1 + 5 runs")
    da.ghost_fill(ctx, W.grid, W.phi, W.ghost_modes, W.ghost_values)
    da.tvd_explicit_step(ctx, W.grid, da.VANLEER, W.phi, W.vel, W.dt, W.phi_work)
    torch.cuda.synchronize()
    phi = W.phi_work.cpu().numpy()
    go = orc.grid([args.n] * 3, ghost=2)
    t = orc.terms(dt=W.dt, diff=W.D, bc_type=[1] * 6, bc_value=[0.0] * 6)

    def one_step():
        t0 = time.perf_counter()
        rp, col, val, rhs, x0 = orc.assemble(go, t, phi)
        t1 = time.perf_counter()
        x, ok, it, res, _ = orc.bicgstab2(rp, col, val, rhs, x0, TOL * float(np.linalg.norm(rhs)), 1000)
        t2 = time.perf_counter()
        return t2 - t0, t1 - t0, t2 - t1, it, ok

    orc.set_threads(threads)
    one_step()                                   # warm-up
    runs = [one_step() for _ in range(5)]
    tot = sorted(r[0] for r in runs)
    med = runs[[r[0] for r in runs].index(tot[2])]
    orc.set_threads(1)
    log("[bench] cpu baseline: one thread, 3 runs")
    single = sorted(one_step()[0] for _ in range(3))
    orc.set_threads(threads)
    return {"value": round(1e3 * tot[2], 1), "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"one full {args.n}^3 implicit step (the median of 5 after 1 warm-up): assembly "
                      f"{1e3 * med[1]:.0f} ms + BiCGStab2 {1e3 * med[2]:.0f} ms ({med[3]} passes, converged={med[4]})",
            "runs_ms": [round(1e3 * v, 1) for v in tot],
            "spread_ms": round(1e3 * (tot[-1] - tot[0]), 1),
            "single_thread_ms": round(1e3 * single[1], 1),
            "single_thread_runs_ms": [round(1e3 * v, 1) for v in single],
            "host": _host_info()}


if __name__ == "__main__":
    main()
