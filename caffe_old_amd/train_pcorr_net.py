#!/usr/bin/env python3
"""Train the learned pressure-correction preconditioner on the GPU.

Training pairs come from the plain solver: seeded 3D lid-driven cavities on an n^3
uniform grid with lid speed U ~ U(0.5, 2) and Reynolds number Re ~ logU(100, 2000)
(mu = U / Re, rho = 1, unit cube), PISO with 2 pressure correctors, dt = 0.5 h / U
(CFL 0.5).  After every pressure-correction solve (to true FP64 relative residual 1e-8)
the p' system -- FP32 matrix, right-hand side scaled to unit norm, converged correction
-- is appended to the device-resident training set.  Half of the batch entries (by
default) are exact synthetic pairs on the stored matrices (x = random Neumann cosine
modes of one frequency band, b = A x): the flow's right-hand sides excite only some
bands, while a preconditioner meets every band in the Krylov residuals, and a network
trained on the flow pairs alone stalls flexible CG.  Training minimises the relative
energy error of one learned two-grid cycle (Adam, cosine-decayed rate, all on the GPU
through the library's own forward and backward kernels).  Held-out cavities (seeds from
10000) measure the loss after training.

    python -m caffe_old_amd.train_pcorr_net --n 128 --runs 80 --steps 25 --pairs 2000 \
        --train-steps 24000 --select --synthetic 0.75 --lr 5e-4

(the command that made the shipped 128^3 weights)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from caffe_old_amd import _lib as L  # noqa: E402
from caffe_old_amd import case as cs  # noqa: E402
from caffe_old_amd.learned import PressureNet, trained_weights_path  # noqa: E402
from caffe_old_amd.solver import make_solver  # noqa: E402


def cavity_params(seed: int):
    """Lid speed and Reynolds number of training / evaluation cavity `seed`."""
    r = np.random.default_rng(seed)
    U = float(r.uniform(0.5, 2.0))
    Re = float(math.exp(r.uniform(math.log(100.0), math.log(2000.0))))
    return U, Re


def cavity(n: int, seed: int, steps: int):
    U, Re = cavity_params(seed)
    h = 1.0 / n
    return cs.lid_cavity(n, U / Re, U, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.5 * h / U,
                         n_pcorr=2, n_inner=1, max_iterations=steps + 1), U, Re


def piso_step_collect(s, net, mode=L.SOLVE_IR_FP32, collect=True, reports=None, skipped=None):
    """One PISO step through the sub-step API (the order of Piso::solve, Piso.cpp:12-29),
    collecting the p' system after each pressure-correction solve.  A system that the
    library refuses as a training pair (zero energy norm x^T A x) is left out and noted in
    `skipped` with its solve report."""
    dt = s.case.time_step
    s.save_u0()
    s.momentum_assemble(dt)
    s.momentum_solve()
    s.rhie_chow()
    for _ in range(s.case.n_pcorr):
        s.pcorr_assemble()
        r = s.pcorr_solve(mode=mode, rtol=1e-8, maxit=100000, inner_rtol=1e-2 if mode == L.SOLVE_IR_FP32 else 1e-4)
        if reports is not None:
            reports.append(r)
        if collect and net is not None:
            try:
                net.collect(s)
            except L.CaffeError as e:
                if skipped is None:
                    raise
                skipped.append((str(e), r))
        s.correct()


MAX_COURANT = 0.75


def generate(net, n, seeds, steps, log, pairs=None, max_courant=MAX_COURANT):
    """Collect the p' systems of `steps` PISO steps of each seeded cavity, stopping once
    the training set holds `pairs` systems (None: no limit).

    With alpha_u = alpha_p = 1 the reference's PISO correctors (Piso.cpp:12-29, no
    neighbour-correction term) go unstable on the fine grid at low Reynolds number, where
    the viscous coefficients outweigh the transient one (6 mu dt / (rho h^2) = 3 / (Re h)
    at CFL 0.5 is about 1.9 for Re 203 at 128^3): the velocity passes the lid speed within
    about ten steps and grows without bound.  A cavity is therefore left at the first step
    that starts from a Courant number above `max_courant` (the reference's
    dt (|u|/dx + |v|/dy + |w|/dz), Simple.cpp:382-408; 0.75 at CFL 0.5 is |u| + |v| + |w| at
    1.5 times the lid speed), so that its systems come from bounded flow only."""
    t0 = time.perf_counter()
    for sd in seeds:
        if pairs is not None and net.dataset_size >= pairs:
            break
        c, U, Re = cavity(n, sd, steps)
        s = make_solver(c)
        s.set_solve_opts("momentum", rtol=1e-8, maxit=5000)
        for k in range(steps):
            if k > 0 and max_courant is not None:
                co = s.step_report()["courant"]
                if not co <= max_courant:
                    log(f"  cavity seed {sd}: left at step {k}, Courant number {co:.3g} > {max_courant}")
                    break
            skipped = []
            piso_step_collect(s, net, skipped=skipped)
            for msg, r in skipped:
                log(f"  cavity seed {sd} step {k}: pair left out ({msg}); p' solve: its {r.get('iters')} "
                    f"rel {r.get('relres')} converged {r.get('converged')}")
            if pairs is not None and net.dataset_size >= pairs:
                break
        s.close()
        log(f"  cavity seed {sd}: U {U:.3f} Re {Re:.0f} -> {net.dataset_size} pairs ({time.perf_counter() - t0:.1f} s)")


def validation_systems(n, seeds, step, log, max_courant=MAX_COURANT, correctors=1):
    """Solver contexts holding the p' systems of the first `correctors` pressure
    correctors of PISO step `step` of each seeded cavity (seeds apart from the training and
    held-out ones), one context per system; a cavity whose flow is no longer bounded by
    then (Courant number above max_courant) is passed over."""
    out = []
    for sd in seeds:
        for corr in range(correctors):
            c, U, Re = cavity(n, sd, step + 1)
            s = make_solver(c)
            s.set_solve_opts("momentum", rtol=1e-8, maxit=5000)
            for _ in range(step):
                piso_step_collect(s, None, collect=False)
            co = s.step_report()["courant"]
            if not co <= max_courant:
                log(f"  validation cavity seed {sd}: passed over (Courant number {co:.3g})")
                s.close()
                break
            s.save_u0()
            s.momentum_assemble(s.case.time_step)
            s.momentum_solve()
            s.rhie_chow()
            for _ in range(corr):   # the correctors before this one, as Piso::solve runs them
                s.pcorr_assemble()
                s.pcorr_solve(mode=L.SOLVE_IR_FP32, rtol=1e-8, maxit=100000, inner_rtol=1e-2)
                s.correct()
            s.pcorr_assemble()
            log(f"  validation cavity seed {sd}: U {U:.3f} Re {Re:.0f}, p' system of step {step}, corrector {corr + 1}")
            out.append(s)
    return out


def validate(net, systems, maxit=3000):
    """Flexible-CG iterations of the learned solve (to a true FP64 relative residual of
    1e-8, FP32 inner iterations restarted every 1e-4) summed over the validation systems;
    a solve that does not converge, or falls back, counts maxit."""
    total = 0
    for s in systems:
        s.attach_net(net)
        r = s.pcorr_solve(mode=L.SOLVE_LEARNED_FCG, rtol=1e-8, maxit=maxit, inner_rtol=1e-4)
        s.attach_net(None)
        total += r["iters"] if (r["converged"] and not r["fell_back"]) else maxit
    return total


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--runs", type=int, default=16)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--pairs", type=int, default=None,
                    help="stop collecting once the training set holds this many systems (default: runs x steps x 2)")
    ap.add_argument("--eval-runs", type=int, default=4)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--synthetic", type=float, default=0.5,
                    help="probability that a batch entry is a synthetic exact pair (cosine modes, b = A x)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--select", action="store_true",
                    help="keep the parameters (checked after every twentieth of the training) with the fewest "
                         "flexible-CG iterations on validation cavities (seeds 20000+)")
    ap.add_argument("--val-runs", type=int, default=3)
    ap.add_argument("--val-step", type=int, default=6)
    ap.add_argument("--val-correctors", type=int, default=1, choices=[1, 2],
                    help="validation systems from the first corrector only, or from both")
    ap.add_argument("--post", type=int, default=1,
                    help="post-smoothing sweeps of the cycle the loss is taken on (the flexible CG applies 3)")
    a = ap.parse_args(argv)

    def log(*x):
        print(*x, flush=True)

    net = PressureNet(seed=a.seed)
    net.set_post_sweeps(a.post)
    log(f"generating training pairs: up to {a.runs} cavities x {a.steps} PISO steps x 2 correctors at {a.n}^3"
        + (f", {a.pairs} pairs" if a.pairs else ""))
    generate(net, a.n, range(a.runs), a.steps, log, a.pairs)
    n_pairs = net.dataset_size
    loss0 = float(net.eval().mean())
    log(f"mean loss before training {loss0:.4f}")
    val, val_curve, best = [], [], None
    if a.select:
        log(f"validation systems: {a.val_runs} cavities (seeds 20000+)")
        val = validation_systems(a.n, range(20000, 20000 + a.val_runs), a.val_step, log, correctors=a.val_correctors)
    t0 = time.perf_counter()
    losses = []
    chunk = max(1, a.train_steps // 20)
    done = 0
    while done < a.train_steps:
        k = min(chunk, a.train_steps - done)
        # one call per chunk keeps a progress line per chunk; the cosine schedule runs
        # over the whole training (lr of this chunk = schedule at its start)
        lr = a.lr * 0.5 * (1.0 + math.cos(math.pi * done / a.train_steps))
        part = net.train(k, a.batch, lr, seed=a.seed * 1000003 + done, synthetic=a.synthetic)
        losses.extend(part.tolist())
        done += k
        vtxt = ""
        if val:
            its = validate(net, val)
            val_curve.append([done, its])
            if best is None or its < best[1]:
                best = (done, its, net.get_params())
            vtxt = f"  validation FCG iterations {its}"
        log(f"  step {done}/{a.train_steps}: batch loss {np.mean(part[-max(1, k // 4):]):.4f}  lr {lr:.2e}  "
            f"({time.perf_counter() - t0:.1f} s){vtxt}")
    train_s = time.perf_counter() - t0
    if best is not None:
        net.set_params(best[2])
        log(f"selected the parameters after step {best[0]}: {best[1]} validation FCG iterations")
    for s_ in val:
        s_.close()
    loss_train = float(net.eval().mean())
    log(f"mean training-set loss after {a.train_steps} steps: {loss_train:.4f} ({train_s:.1f} s)")

    net.clear()
    log(f"generating held-out pairs: {a.eval_runs} cavities (seeds 10000+)")
    generate(net, a.n, range(10000, 10000 + a.eval_runs), a.steps, log)
    ev = net.eval()
    log(f"held-out loss: mean {ev.mean():.4f} max {ev.max():.4f}")
    net.clear()

    out = a.out or trained_weights_path(a.n)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    meta = {"grid": a.n, "runs": a.runs, "steps_per_run": a.steps, "pairs": n_pairs,
            "train_steps": a.train_steps, "batch": a.batch, "lr": a.lr, "seed": a.seed, "synthetic": a.synthetic,
            "post_sweeps": a.post, "val_correctors": a.val_correctors,
            "loss_before": loss0, "loss_train_after": loss_train,
            "loss_heldout_mean": float(ev.mean()), "loss_heldout_max": float(ev.max()),
            "train_seconds": train_s, "selected_step": best[0] if best else None,
            "validation_fcg_iterations": val_curve, "loss_curve_every_50": [float(np.mean(losses[i:i + 50])) for i in range(0, len(losses), 50)]}
    net.save(out, meta)
    log(f"saved {out}")
    net.close()


if __name__ == "__main__":
    main()
