"""Training and evaluation of the learned warm start on solver trajectories.

    python -m dare_amd.warmstart train --n 128 --train-seeds 200 --steps 200 --out ws.npz
    python -m dare_amd.warmstart eval  --n 128 --weights ws.npz --eval-seeds 20
    python -m dare_amd.warmstart train --n 256 --workload vardiff --train-seeds 400 --steps 6 \
        --crop 64 --lr 1e-3 --lr-final 5e-5 --weights dare_amd/weights/ws_vardiff_128.npz \
        --out dare_amd/weights/ws_vardiff_256.npz        (fine-tuning at 256^3)
    python -m dare_amd.warmstart train --n 128 --coarsen 2 --crop 64 ...   (half resolution)

train: for every training seed, the plain FP64 BiCGStab2 advances the advection-diffusion
trajectory (C5AdvDiff) from the reference's guess (phi^n); each step's (phi^n, phi^{n-1},
rhs) -> solution pair is one online Adam step of the network (hand-written BF16 MFMA forward
and backward kernels). Nothing is stored between steps.

eval: on held-out seeds, every step is solved from four initial guesses — zero, phi^n (the
reference's choice), the linear extrapolation 2 phi^n - phi^{n-1}, and the network's
correction of that extrapolation — to the same FP64 tolerance; passes, the initial residual
and wall time (the guess's cost included) are reported per guess.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import Context, bicgstab2, residual
from .learned import WarmStartNet
from .workloads import WORKLOADS

TOL = 1e-8


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def train(ctx, n=128, train_seeds=range(1000, 1200), steps=200, lr=1e-3, lr_final=1e-4, net=None,
          train_every=1, log_every=60.0, max_seconds=None, seed=0, workload="advdiff", crop=None, coarsen=1):
    """Online training on plain-solver trajectories. With `crop` = c, each step trains on one
    random c^3 sub-box of the grid (the network is fully convolutional; with coarsening 2 the
    crop origin is even, so the sub-box's 2x2x2 blocks are blocks of the grid). Returns
    (net, info)."""
    import torch
    make = WORKLOADS[workload]
    crng = np.random.default_rng(seed + 17)
    net = net or WarmStartNet(ctx, [n, n, n], in_channels=make.in_channels_for(), seed=seed, coarsen=coarsen)
    cf = net.coarsen
    seeds = list(train_seeds)
    total = len(seeds) * max(steps - 1, 1) // train_every
    done, t_start, t_log = 0, time.perf_counter(), time.perf_counter()
    A = None
    losses, passes = [], []
    for si, s in enumerate(seeds):
        W = make(ctx, s, n=n, A=A)
        A = W.A
        for k in range(steps):
            W.assemble()
            r = bicgstab2(ctx, A, W.rhs, W.x, tol=TOL, tol_kind=1, max_it=1000)
            passes.append(r["iterations"])
            if k >= 1 and k % train_every == 0:
                frac = min(done / max(total, 1), 1.0)
                cur_lr = lr * (lr_final / lr) ** frac
                box = None
                if crop:
                    o = [cf * int(crng.integers(0, (n - crop) // cf + 1)) for _ in range(3)]
                    box = o + [crop] * 3
                losses.append(net.train_step(W.grid, W.phi, W.phi_prev, W.xb, W.r0, W.rhs_scale, W.x, lr=cur_lr,
                                             extras=W.extras, crop=box))
                done += 1
            W.advance(W.x)
            now = time.perf_counter()
            if now - t_log > log_every:
                t_log = now
                _log(f"[train] seed {si + 1}/{len(seeds)} step {k} samples {done} "
                     f"loss(last 50) {np.mean(losses[-50:]):.4e} elapsed {now - t_start:.0f}s")
            if max_seconds and now - t_start > max_seconds:
                break
        if max_seconds and time.perf_counter() - t_start > max_seconds:
            break
    torch.cuda.synchronize()
    info = {"workload": workload, "samples": done, "seeds_used": si + 1, "steps": steps, "n": n, "crop": crop,
            "coarsen": cf,
            "seconds": round(time.perf_counter() - t_start, 1),
            "loss_first_100": float(np.mean(losses[:100])) if losses else None,
            "loss_last_100": float(np.mean(losses[-100:])) if losses else None,
            "plain_passes_mean": float(np.mean(passes)) if passes else None}
    return net, info


def evaluate(ctx, net, n=128, eval_seeds=range(5000, 5020), steps=200, tol=TOL, log_every=60.0,
             workload="advdiff"):
    """Passes and time per step for each initial guess on held-out trajectories."""
    import torch
    make = WORKLOADS[workload]
    modes = ["zero", "previous", "extrapolated", "learned"]
    stats = {m: {"passes": [], "solve_s": [], "guess_s": [], "r0": [], "conv": []} for m in modes}
    A = None
    xg = None
    t_log = time.perf_counter()
    for si, s in enumerate(eval_seeds):
        W = make(ctx, s, n=n, A=A)
        A = W.A
        if xg is None:
            xg = torch.empty_like(W.x)
            rbuf = torch.empty_like(W.x)
        for k in range(steps):
            W.assemble(base=False)
            x_prev = W.x.clone()                    # phi^n in solver order (fill wrote x0 = phi^n)
            keep = None
            for m in modes:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                # each guess is timed with everything it needs: the extrapolation xb, and for the
                # network also the residual r0 = b - A xb it reads
                if m == "zero":
                    xg.zero_()
                elif m == "previous":
                    xg.copy_(x_prev)
                elif m == "extrapolated":
                    W.base_guess(with_residual=False, out=xg)   # 2 phi^n - phi^{n-1}
                else:
                    W.base_guess()
                    net.predict(W.grid, W.phi, W.phi_prev, W.xb, W.r0, W.rhs_scale, xg, extras=W.extras)
                torch.cuda.synchronize()
                guess_s = time.perf_counter() - t0
                if k >= 1:                          # initial residual of the guess (not timed)
                    rn = residual(ctx, A, W.rhs, xg, rbuf, norm=True)
                    stats[m]["r0"].append(rn / float(torch.linalg.norm(W.rhs)))
                    torch.cuda.synchronize()
                t1 = time.perf_counter()
                r = bicgstab2(ctx, A, W.rhs, xg, tol=tol, tol_kind=1, max_it=1000)
                torch.cuda.synchronize()
                t2 = time.perf_counter()
                if k >= 1:                          # step 0 has no phi^{n-1}: all guesses coincide
                    stats[m]["passes"].append(r["iterations"])
                    stats[m]["conv"].append(r["converged"])
                    stats[m]["solve_s"].append(t2 - t1)
                    stats[m]["guess_s"].append(guess_s)
                if m == "previous":
                    keep = xg.clone()
            W.advance(keep)
            if time.perf_counter() - t_log > log_every:
                t_log = time.perf_counter()
                _log(f"[eval] seed {si + 1}/{len(eval_seeds)} step {k}: "
                     + ", ".join(f"{m} {np.mean(stats[m]['passes'] or [0]):.2f}" for m in modes))
    out = {}
    for m in modes:
        p = np.array(stats[m]["passes"], float)
        out[m] = {"passes_mean": float(p.mean()), "passes_total": int(p.sum()),
                  "solve_ms_mean": 1e3 * float(np.mean(stats[m]["solve_s"])),
                  "guess_ms_mean": 1e3 * float(np.mean(stats[m]["guess_s"])),
                  "step_ms_mean": 1e3 * float(np.mean(stats[m]["solve_s"]) + np.mean(stats[m]["guess_s"])),
                  "initial_rel_residual_geomean": float(np.exp(np.mean(np.log(np.maximum(stats[m]["r0"], 1e-300))))),
                  "all_converged": bool(all(stats[m]["conv"]))}
    base = out["previous"]
    for m in modes:
        out[m]["passes_saved_vs_previous"] = 1.0 - out[m]["passes_mean"] / base["passes_mean"]
        out[m]["time_saved_vs_previous"] = 1.0 - out[m]["step_ms_mean"] / base["step_ms_mean"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval", "train-eval"])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="advdiff")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--train-seeds", type=int, default=200)
    ap.add_argument("--eval-seeds", type=int, default=20)
    ap.add_argument("--eval-steps", type=int, default=None)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-4)
    ap.add_argument("--train-every", type=int, default=1)
    ap.add_argument("--crop", type=int, default=None)
    ap.add_argument("--coarsen", type=int, default=None, help="1 or 2 (default: the weights' own, else 1)")
    ap.add_argument("--max-seconds", type=float, default=None)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    ctx = Context(0)
    report = {}
    net = None
    if a.cmd in ("train", "train-eval"):
        init, cf = None, a.coarsen or 1
        if a.weights:
            init, meta = WarmStartNet.load_params(a.weights)
            cf = a.coarsen or meta["coarsen"]
        net = WarmStartNet(ctx, [a.n] * 3, in_channels=WORKLOADS[a.workload].in_channels_for(), params=init,
                           coarsen=cf)
        net, info = train(ctx, n=a.n, train_seeds=range(1000, 1000 + a.train_seeds), steps=a.steps, lr=a.lr,
                          lr_final=a.lr_final, net=net, train_every=a.train_every, max_seconds=a.max_seconds, workload=a.workload,
                          crop=a.crop)
        report["train"] = info
        _log(f"[train] {info}")
        if a.out:
            net.save(a.out, meta={"train": info, "workload": a.workload, "tol": TOL})
    if a.cmd in ("eval", "train-eval"):
        if net is None:
            net = WarmStartNet.load(ctx, a.weights, [a.n] * 3)
        ev = evaluate(ctx, net, n=a.n, eval_seeds=range(5000, 5000 + a.eval_seeds), steps=a.eval_steps or a.steps,
                      workload=a.workload)
        report["eval"] = ev
        _log(f"[eval] {json.dumps(ev)}")
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps(report))
    return 0


if __name__ == "__main__":
    sys.exit(main())
