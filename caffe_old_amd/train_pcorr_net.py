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

    python -m caffe_old_amd.train_pcorr_net --n 128 --runs 16 --steps 12 --train-steps 3000
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


def piso_step_collect(s, net, mode=L.SOLVE_IR_FP32, collect=True, reports=None):
    """One PISO step through the sub-step API (the order of Piso::solve, Piso.cpp:12-29),
    collecting the p' system after each pressure-correction solve."""
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
            net.collect(s)
        s.correct()


def generate(net, n, seeds, steps, log):
    t0 = time.perf_counter()
    for sd in seeds:
        c, U, Re = cavity(n, sd, steps)
        s = make_solver(c)
        s.set_solve_opts("momentum", rtol=1e-8, maxit=5000)
        for _ in range(steps):
            piso_step_collect(s, net)
        s.close()
        log(f"  cavity seed {sd}: U {U:.3f} Re {Re:.0f} -> {net.dataset_size} pairs ({time.perf_counter() - t0:.1f} s)")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--runs", type=int, default=16)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--eval-runs", type=int, default=4)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--synthetic", type=float, default=0.5,
                    help="probability that a batch entry is a synthetic exact pair (cosine modes, b = A x)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)

    def log(*x):
        print(*x, flush=True)

    net = PressureNet(seed=a.seed)
    log(f"generating training pairs: {a.runs} cavities x {a.steps} PISO steps x 2 correctors at {a.n}^3")
    generate(net, a.n, range(a.runs), a.steps, log)
    loss0 = float(net.eval().mean())
    log(f"mean loss before training {loss0:.4f}")
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
        log(f"  step {done}/{a.train_steps}: batch loss {np.mean(part[-max(1, k // 4):]):.4f}  lr {lr:.2e}  "
            f"({time.perf_counter() - t0:.1f} s)")
    train_s = time.perf_counter() - t0
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
    meta = {"grid": a.n, "runs": a.runs, "steps_per_run": a.steps, "pairs": a.runs * a.steps * 2,
            "train_steps": a.train_steps, "batch": a.batch, "lr": a.lr, "seed": a.seed, "synthetic": a.synthetic,
            "loss_before": loss0, "loss_train_after": loss_train,
            "loss_heldout_mean": float(ev.mean()), "loss_heldout_max": float(ev.max()),
            "train_seconds": train_s, "loss_curve_every_50": [float(np.mean(losses[i:i + 50])) for i in range(0, len(losses), 50)]}
    net.save(out, meta)
    log(f"saved {out}")
    net.close()


if __name__ == "__main__":
    main()
