#!/usr/bin/env python3
"""Benchmark: PISO time steps per second on the 128^3 lid-driven cavity (Re 1000, two
pressure correctors, dt = 0.5 h / U), plus the pressure-correction solve time to a true
FP64 relative residual of 1e-8: learned flexible CG (BF16 network inside FP32 iterations,
FP64 refinement) and FP32-inner / FP64-refinement CG vs plain FP64 CG, same system.

One "step" = Piso::solve(dt) (Piso.cpp:12-29): momentum assembly + 3 BiCGStab solves,
Rhie-Chow, then 2 x (p' assembly + p' Krylov solve + correct), continuity and Courant.
Inputs are resident in HBM before the timed region; the timed region contains K full
steps bracketed by stream synchronisation.  The p' solves of the timed steps use the
learned FCG when trained weights for the grid ship with the package (--mode auto).

roofline: the kernel with the largest total time in the rocprof trace of the timed steps
(k_f_k2pre, the learned FCG's update and pre-smoothing kernel; k_cg_k1 for --mode ir/plain/
sgs; the multigrid's fine post-smoothing sweep for --mode mg), its duration averaged over
its launches in the timed steps from HIP event pairs recorded around them on the solver's
stream, its algorithmic bytes per launch (interior cells x bytes per cell), and the HBM
bytes per launch counted by rocprofv3 (profiles/traffic.json, calibrated against a copy
kernel).

Beside the contract's K timed steps the line reports the configuration as stated, 100 steps
from rest timed as a whole (config.full_run_100_steps), the p' solve of one system by every
method (the red-black SGS-preconditioned CG being the judge as strong as the reference's
SOR), the CPU baseline on all host cores over the same part of the run, and the error of
every field against the FP64 oracle after each of four steps.

Prints one JSON line.  --gpus N > 1 is not supported: the whole domain lives on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _cpu_threads():
    """every host core this process may run on (the box's CPU share), or OMP_NUM_THREADS when
    that is set lower"""
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_baseline(n, steps=2, warmup=5):
    """The oracle (C++ FP64 restatement of the reference path, BiCGStab + block-local SSOR
    at rtol 1e-5 like the reference's PETSc settings, OpenMP threads standing in for MPI
    ranks, one SSOR block per thread) on the same 128^3 case and the same part of the run as
    the GPU's timed steps: `warmup` untimed PISO steps from rest, then `steps` timed ones."""
    from caffe_old_amd import case as cs
    from oracle.pyoracle import KSP_PETSC_LIKE, Oracle
    th = _cpu_threads()
    c = cs.cavity_piso_128(n, steps=steps + warmup)
    o = Oracle(c, ksp=KSP_PETSC_LIKE, rtol=1e-5, maxit=10000, n_blocks=th, n_threads=th)
    for _ in range(warmup):
        o.solve()
    t0 = time.perf_counter()
    for _ in range(steps):
        o.solve()
    dt = time.perf_counter() - t0
    return {"value": steps / dt, "unit": "steps/s", "cores": th, "kind": "port", "cpu_model": _cpu_model(),
            "sample": f"steps {warmup + 1}-{warmup + steps} of the {n}^3 cavity PISO run from rest ({warmup} untimed), "
                      f"oracle BiCGStab+SSOR rtol 1e-5, {th} OpenMP threads on {_cpu_model()}, {dt:.1f} s timed"}


def field_errors(n, mode, inner, net, mom_mode, steps=4, post=None):
    """Part of the CPU leg, the oracle as the checker: the error of every field against the
    FP64 oracle on this configuration.  A fresh GPU run with the timed run's solve settings
    (true relres 1e-8) and the oracle with tight solves (true relres 1e-12) each take `steps`
    PISO steps from rest; relative L2 over the interior cells, p after removing the mean (the
    pure-Neumann pressure is defined up to a constant)."""
    import numpy as np
    from caffe_old_amd import case as cs
    from caffe_old_amd.solver import make_solver
    from oracle.pyoracle import Oracle, interior
    th = _cpu_threads()
    c = cs.cavity_piso_128(n, steps=steps)
    g = make_solver(c)
    g.set_solve_opts("momentum", mode=mom_mode, rtol=1e-8, maxit=5000)
    g.set_solve_opts("pcorr", mode=mode, rtol=1e-8, maxit=100000, inner_rtol=inner)
    if net is not None:
        g.attach_net(net)
    if post is not None:
        g.set_fcg_post_sweeps(post)
    o = Oracle(c, rtol=1e-12, maxit=20000, n_blocks=th, n_threads=th)
    t0 = time.perf_counter()

    def rel(x, y):
        ny = float(np.linalg.norm(y))
        if not np.isfinite(ny):
            return float("nan")
        return float(np.linalg.norm(x - y)) / ny if ny > 0 else float(np.linalg.norm(x))

    sh = g.shape
    per_step = []
    for _ in range(steps):
        g.solve()
        o.solve()
        ug, uo = interior(g.get("u"), sh), interior(o.get("u"), sh)
        pg, po = interior(g.get("p"), sh), interior(o.get("p"), sh)
        e = {comp: rel(ug[i], uo[i]) for i, comp in enumerate("uvw")}
        e["p"] = rel(pg - pg.mean(), po - po.mean())
        per_step.append(e)
    out = dict(per_step[-1])
    g.close()
    out.update({"steps": steps, "max_over_steps": max(max(e.values()) for e in per_step),
                "per_step_max": [max(e.values()) for e in per_step], "gpu_rtol": 1e-8, "oracle_rtol": 1e-12,
                "norm": "relative L2 over interior cells, p with the mean removed",
                "seconds": round(time.perf_counter() - t0, 1)})
    return out


def full_run(n, mode, inner, net, mom_mode, post=None):
    """The configuration as stated: 100 PISO steps of the n^3 cavity from rest, in a fresh
    context with the timed run's solve settings, timed as a whole (synchronised before and
    after).  Reported beside the headline's K timed steps."""
    from caffe_old_amd import case as cs
    from caffe_old_amd.solver import make_solver
    c = cs.cavity_piso_128(n, steps=100)
    g = make_solver(c)
    g.set_solve_opts("momentum", mode=mom_mode, rtol=1e-8, maxit=5000)
    g.set_solve_opts("pcorr", mode=mode, rtol=1e-8, maxit=100000, inner_rtol=inner)
    if net is not None:
        g.attach_net(net)
    if post is not None:
        g.set_fcg_post_sweeps(post)
    g.sync()
    t0 = time.perf_counter()
    fell = not_conv = 0
    for _ in range(c.n_solves):
        g.solve()
        m, p = g.reports()
        fell += p["fell_back"]
        not_conv += 0 if (p["converged"] and m["converged"]) else 1
    g.sync()
    dt = time.perf_counter() - t0
    cont = g.continuityError
    courant = g.step_report()["courant"]
    import numpy as np
    ni, nj, nk = g.shape
    u = g.get("u").reshape(3, nk + 2, nj + 2, ni + 2)[:, 1:-1, 1:-1, 1:-1]
    umax = float(np.abs(u).max())   # interior cells; the lid speed is 1
    g.close()
    return {"steps": c.n_solves, "seconds": round(dt, 3), "steps_per_s": round(c.n_solves / dt, 3),
            "p_fell_back_steps": fell, "not_converged_steps": not_conv, "final_continuity": cont,
            "final_courant": round(courant, 4), "final_max_abs_u_over_lid": round(umax, 4),
            "note": "from rest, including the first solves that allocate work and capture graphs"}


def coupled_bench(a):
    """The Coupled solver (Coupled.cpp:38-338) on the n^3 lid-driven cavity (Re 100, steady,
    relaxationFactorMomentum 0.7): one "step" is Coupled::solve, i.e. assembly of the
    4-variable system, its BiCGStab solve to a true FP64 relative residual of 1e-8 from
    x = 0 (plus the null-vector gauge of the closed cavity), and the field update.  The
    BiCGStab is right-preconditioned by the block lower triangle with a multigrid V-cycle on
    the pressure block (--mode auto or mg), or by Jacobi (--mode plain).  The roofline names
    the cycle's level-0 post-smoothing sweep, or KA of the Jacobi BiCGStab (k_cp_a: 56
    coefficient planes and the four variables' operands per cell).  The CPU baseline is the oracle with the reference's
    PETSc-like settings (BiCGStab + block SSOR, rtol 1e-5 on the preconditioned norm)."""
    from caffe_old_amd import case as cs
    from caffe_old_amd.solver import make_solver
    from caffe_old_amd import _lib as L
    n = a.coupled
    mg = a.mode in ("mg", "auto")
    probe = "mg_post0" if mg else "coupled_bicg_a"
    case = cs.lid_cavity(n, 0.01, algo="coupled", max_iterations=a.steps + a.warmup + 1)
    s = make_solver(case)
    s.set_solve_opts("pcorr", mode=L.SOLVE_MG if mg else L.SOLVE_PLAIN_FP64, rtol=1e-8, maxit=200000)
    for _ in range(a.warmup):
        s.solve()
    s.kernel_probe(probe, True)
    s.sync()
    t0 = time.perf_counter()
    its, steps_detail = [], []
    for _ in range(a.steps):
        s.solve()
        p = s.reports()[1]
        its.append(p["iters"])
        steps_detail.append({"iters": p["iters"], "converged": p["converged"], "relres": p["relres"],
                             "fell_back": p["fell_back"]})
    s.sync()
    elapsed = time.perf_counter() - t0
    kms, klaunches, kbytes = s.kernel_probe_read()
    s.kernel_probe(probe, False)
    bad = sum(1 for d in steps_detail if not d["converged"] or d["fell_back"])
    if bad:
        print(f"bench.py --coupled: {bad} of {a.steps} timed steps did not converge or fell back", file=sys.stderr)
    if a.calibrate_copy:
        s.time_kernel("copy_fp64", reps=5)
    ms_per_step = 1000.0 * elapsed / a.steps
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(f"{probe}_coupled_{n}" if mg else f"coupled_bicg_a_{n}")
    out = {
        "metric": f"Coupled iterations/s, {n}^3 lid-driven cavity",
        "value": round(1000.0 / ms_per_step, 4), "unit": "iterations/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic: lid-driven cavity from rest (no external data)",
        "config": {"workload": f"3D lid-driven cavity {n}^3 uniform hex, Re 100 (mu 1e-2, lid 1), Coupled solver, steady, "
                               f"relaxationFactorMomentum 0.7, 4-variable system solved to true relres 1e-8",
                   "cells": case.n_cells, "bicgstab_iters_per_step": its, "steps_detail": steps_detail,
                   "all_steps_converged_without_fallback": bad == 0,
                   "preconditioner": "block triangle, multigrid V-cycle on the pressure block" if mg else "Jacobi"},
        "roofline": {"bound": "hbm", "kernel": probe, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "ms_per_launch": round(kms, 5), "bytes_per_launch": kbytes, "launches_timed": klaunches},
    }
    if not a.no_cpu:
        from oracle.pyoracle import KSP_PETSC_LIKE, Oracle
        th = _cpu_threads()
        o = Oracle(case, ksp=KSP_PETSC_LIKE, rtol=1e-5, maxit=10000, n_blocks=th, n_threads=th)
        o.solve()
        t0 = time.perf_counter()
        o.solve()
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "iterations/s", "cores": th, "kind": "port",
                               "sample": f"iteration 2 of the {n}^3 Coupled cavity from rest (1 untimed), oracle "
                                         f"BiCGStab+SSOR rtol 1e-5, {th} OpenMP threads, {dt:.1f} s timed"}
    print(json.dumps(out), flush=True)


def workload_bench(a):
    """The two full-size configurations beside the headline one, each on one GPU:
      ib-channel  channel 4 x 1 x 1 on 256 x 64 x 64 cells with an immersed sphere (D 0.2 at
                  (1, .5, .5)), uniform inlet u = 1, Re_D 200, PISO 2 correctors, dt 0.5/64;
      skewed      256^3 cavity whose interior nodes move by up to 0.15 h (seeded Philox),
                  Re 1000, PISO 2 correctors, dt 0.5/256.
    One step is Piso::solve / IbPiso::solve.  The p' solves of the timed steps use --mode
    (auto: the learned FCG on the skewed cavity, as that configuration names it; the
    multigrid BiCGStab on the IB channel's ghost-row systems, which the learned cycle does
    not take).  After the timed steps the same p' system is solved by each method to a true
    FP64 relres of 1e-8.  The roofline names the dominant p' kernel of the timed steps;
    `device_bytes` is the context's HBM footprint."""
    from caffe_old_amd import _lib as L
    from caffe_old_amd import case as cs
    from caffe_old_amd.learned import PressureNet, trained_weights_path
    from caffe_old_amd.solver import make_solver
    ib = a.workload == "ib-channel"
    if ib:
        case = cs.channel_ib_sphere(256, 64, 64, steps=a.steps + a.warmup)
        desc = ("channel 4x1x1, 256x64x64 uniform hex, immersed sphere D 0.2 at (1,.5,.5), inlet u 1, Re_D 200, "
                "IbPiso 2 correctors, dt 0.5/64")
        s = make_solver(case)
    else:
        case = cs.skewed_cavity(256, steps=a.steps + a.warmup)
        case.non_orth = a.non_orth
        desc = ("256^3 lid-driven cavity, interior nodes moved by up to 0.15 h (Philox seed 0), Re 1000, PISO 2 "
                "correctors, dt 0.5/256, non-orthogonal term " +
                ("corrected (grad u faces filled)" if a.non_orth else "as the reference (zero)"))
        s = make_solver(case, uniform=False)
    net = None
    wpath = trained_weights_path(128)
    if not ib and os.path.exists(wpath):
        # the network trained on 128^3 cavities, applied on the 128^3 coarse grid of this mesh
        net = PressureNet(seed=0).load(wpath)
        s.attach_net(net)
    # one post-smoothing sweep on the skewed 256^3 cavity (measured there: 10.5 steps/s with 1,
    # 9.5 with 2, 9.0 with 3; the network works on a 128^3 coarse grid it was not trained on,
    # and more smoothing saves few iterations), unless --fcg-post says otherwise
    s.set_fcg_post_sweeps(a.fcg_post if a.fcg_post is not None else 1)
    mode_name = a.mode if a.mode != "auto" else ("mg" if ib or net is None else "learned")
    modes = {"ir": L.SOLVE_IR_FP32, "plain": L.SOLVE_PLAIN_FP64, "learned": L.SOLVE_LEARNED_FCG, "mg": L.SOLVE_MG}
    inner = {"ir": 1e-2, "plain": 1e-2, "learned": 1e-4, "mg": 1e-2}
    s.set_solve_opts("momentum", mode=L.SOLVE_IR_FP32, rtol=1e-8, maxit=5000)
    s.set_solve_opts("pcorr", mode=modes[mode_name], rtol=1e-8, maxit=200000, inner_rtol=inner[mode_name])
    if ib:
        kname = "pcorr_bicg_a_fp64" if mode_name == "plain" else "mg_post0"
    else:
        kname = {"learned": "fcg_k1", "ir": "pcorr_cg_k1_fp32", "plain": "pcorr_cg_k1_fp64", "mg": "mg_post0"}[mode_name]
    for w in range(a.warmup):
        s.solve()
    # the multigrid iterations replay a captured graph, which a recording probe turns off:
    # its timed steps run unprobed and the roofline kernel is timed in steps after them
    probe_in_steps = mode_name != "mg"
    if probe_in_steps:
        s.kernel_probe(kname, True)
    s.sync()
    t0 = time.perf_counter()
    its = []
    for _ in range(a.steps):
        s.solve()
        m, p = s.reports()
        its.append({"p_iters": p["iters"], "fell_back": p["fell_back"], "mom_iters": m["iters"]})
    s.sync()
    elapsed = time.perf_counter() - t0
    if not probe_in_steps:
        s.kernel_probe(kname, True)
        for _ in range(2):
            s.solve()
    kms, klaunches, kbytes = s.kernel_probe_read()
    s.kernel_probe(kname, False)
    if a.calibrate_copy:
        s.time_kernel("copy_fp64", reps=5)
    psolve = {}
    methods = [("plain_fp64", L.SOLVE_PLAIN_FP64, 1e-2), ("ir_fp32", L.SOLVE_IR_FP32, 1e-2), ("mg", L.SOLVE_MG, 1e-2)]
    if net is not None:
        methods.append(("learned_fcg", L.SOLVE_LEARNED_FCG, 1e-4))
    for name, md, inr in methods:   # each method's second solve (the first one also allocates)
        s.pcorr_assemble()
        s.pcorr_solve(mode=md, rtol=1e-8, maxit=200000, inner_rtol=inr)
        r = s.pcorr_solve(mode=md, rtol=1e-8, maxit=200000, inner_rtol=inr)
        psolve[name] = {"ms": round(r["ms"], 3), "iters": r["iters"], "refinements": r["refinements"],
                        "relres": r["relres"], "fell_back": r["fell_back"]}
    ms_per_step = 1000.0 * elapsed / a.steps
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    out = {
        "metric": f"{'IbPiso' if ib else 'PISO'} steps/s, {a.workload}", "value": round(1000.0 / ms_per_step, 4),
        "unit": "steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"plain": "f64", "ir": "f64+f32", "learned": "f64+f32+bf16", "mg": "f64+f32"}[mode_name],
        "data": "synthetic (no external data)",
        "config": {"workload": desc + f", p' solve to true relres 1e-8 ({mode_name})", "cells": case.n_cells,
                   "non_orthogonal_correction": "n/a (uniform mesh)" if ib else ("corrected" if a.non_orth else "reference"),
                   "device_bytes": s.device_bytes(), "steps_detail": its, "p_solve_ms": psolve},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "ms_per_launch": round(kms, 5),
                     "bytes_per_launch": kbytes, "launches_timed": klaunches},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--mode", choices=["auto", "ir", "plain", "learned", "mg", "sgs"], default="auto",
                    help="p' solve: learned FCG (auto when trained weights exist for the grid), FP32 IR, plain FP64, "
                         "multigrid BiCGStab")
    ap.add_argument("--mom-mode", choices=["ir", "plain"], default="ir",
                    help="momentum solves: FP32 BiCGStab inside FP64 refinement, or plain FP64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-full-run", action="store_true", help="skip the 100-step run from rest")
    ap.add_argument("--fcg-post", type=int, default=None, metavar="M",
                    help="damped Jacobi sweeps after the learned cycle's coarse correction (1-4; default the "
                         "library's, 3)")
    ap.add_argument("--non-orth", action="store_true",
                    help="--workload skewed: the corrected mode of the non-orthogonal momentum term")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, 5 launches of a plain copy of 8 P doubles (counter calibration)")
    ap.add_argument("--kernel-only", action="store_true",
                    help="one warm-up step, then only the dominant kernel and the copy kernel (for counter runs)")
    ap.add_argument("--coupled", type=int, default=0, metavar="N",
                    help="measure the Coupled solver on the N^3 cavity instead (its own JSON line)")
    ap.add_argument("--workload", choices=["cavity", "ib-channel", "skewed"], default="cavity",
                    help="ib-channel / skewed: the 256x64x64 IB channel or the skewed 256^3 cavity (own JSON line)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: --gpus {a.gpus} is not supported; this build runs the whole domain on one GPU "
              f"(multi-GPU block decomposition is design-only).", file=sys.stderr)
        sys.exit(2)
    if a.coupled:
        coupled_bench(a)
        return
    if a.workload != "cavity":
        workload_bench(a)
        return

    from caffe_old_amd import _lib as L
    from caffe_old_amd import case as cs
    from caffe_old_amd.solver import make_solver

    from caffe_old_amd.learned import PressureNet, trained_weights_path
    wpath = trained_weights_path(a.n)
    if a.mode == "auto":
        a.mode = "learned" if os.path.exists(wpath) else "ir"
    mode = {"ir": L.SOLVE_IR_FP32, "plain": L.SOLVE_PLAIN_FP64, "learned": L.SOLVE_LEARNED_FCG, "mg": L.SOLVE_MG,
            "sgs": L.SOLVE_SGS_FP64}[a.mode]
    inner = {"ir": 1e-2, "plain": 1e-2, "learned": 1e-4, "mg": 1e-2, "sgs": 1e-2}
    case = cs.cavity_piso_128(a.n, steps=a.steps + a.warmup)
    s = make_solver(case)
    mom_mode = {"ir": L.SOLVE_IR_FP32, "plain": L.SOLVE_PLAIN_FP64}[a.mom_mode]
    s.set_solve_opts("momentum", mode=mom_mode, rtol=1e-8, maxit=5000)
    s.set_solve_opts("pcorr", mode=mode, rtol=1e-8, maxit=100000, inner_rtol=inner[a.mode])
    # the learned preconditioner trained for this grid (caffe_old_amd.train_pcorr_net)
    net = None
    if os.path.exists(wpath):
        net = PressureNet(seed=0).load(wpath)
        s.attach_net(net)
    if a.fcg_post is not None:
        s.set_fcg_post_sweeps(a.fcg_post)
    # the dominant kernel of the timed steps (the largest total time in the rocprof kernel
    # trace of the same command): the learned FCG's update and pre-smoothing kernel, the CG
    # stencil kernel, or the multigrid's fine-level post-smoothing sweep
    kname = {"learned": "fcg_k2pre", "ir": "pcorr_cg_k1_fp32", "plain": "pcorr_cg_k1_fp64", "mg": "mg_post0",
             "sgs": "pcorr_cg_k1_fp64"}[a.mode]

    def log(*x):
        if not a.quiet:
            print(*x, file=sys.stderr, flush=True)

    if a.kernel_only:
        s.solve()
        s.pcorr_assemble()
        for kn in ("pcorr_cg_k1_fp64", "pcorr_cg_k1_fp32", "copy_fp64"):
            ms, by = s.time_kernel(kn, reps=20)
            print(json.dumps({"kernel": kn, "ms": ms, "bytes": by, "GBs": by / ms / 1e6}), flush=True)
        return

    for w in range(a.warmup):
        s.solve()
        log(f"warmup step {w}: continuity {s.continuityError:.3e}")
    # HIP event pairs around that kernel's launches, on its stream; the multigrid iterations
    # replay a captured graph, which a recording probe turns off, so in that mode the kernel
    # is timed in two steps after the timed ones
    probe_in_steps = a.mode != "mg"
    if probe_in_steps:
        s.kernel_probe(kname, True)
    s.sync()
    t0 = time.perf_counter()
    iters_p = 0
    fell_steps = not_conv_steps = 0
    for k in range(a.steps):
        s.solve()
        m, p = s.reports()
        iters_p += p["iters"]
        fell_steps += p["fell_back"]
        not_conv_steps += 0 if (p["converged"] and m["converged"]) else 1
    s.sync()
    elapsed = time.perf_counter() - t0
    if not probe_in_steps:
        s.kernel_probe(kname, True)
        for _ in range(2):
            s.solve()
    kms, klaunches, kbytes = s.kernel_probe_read()
    s.kernel_probe(kname, False)
    if a.calibrate_copy:
        s.time_kernel("copy_fp64", reps=5)
    ms_per_step = 1000.0 * elapsed / a.steps
    rep = s.step_report()
    log(f"timed {a.steps} steps: {ms_per_step:.2f} ms/step, last p' iters {p['iters']}, courant {rep['courant']:.3f}")

    # pressure-correction solve on the same system: plain FP64 (Jacobi CG, and CG with the
    # red-black SGS sweep, the analogue of the reference's PCSOR), FP32-inner / FP64-refinement,
    # multigrid BiCGStab, learned FCG; each method's second solve is reported (the first one of
    # a method in a context also allocates its work and captures its graph)
    psolve = {}
    for name, md, inr in (("plain_fp64", L.SOLVE_PLAIN_FP64, 1e-2), ("sgs_fp64", L.SOLVE_SGS_FP64, 1e-2),
                          ("ir_fp32", L.SOLVE_IR_FP32, 1e-2), ("mg", L.SOLVE_MG, 1e-2),
                          ("learned_fcg", L.SOLVE_LEARNED_FCG, 1e-4)):
        s.pcorr_assemble()
        s.pcorr_solve(mode=md, rtol=1e-8, maxit=100000, inner_rtol=inr)
        r = s.pcorr_solve(mode=md, rtol=1e-8, maxit=100000, inner_rtol=inr)
        psolve[name] = {"ms": round(r["ms"], 3), "iters": r["iters"], "refinements": r["refinements"],
                        "relres": r["relres"], "fell_back": r["fell_back"]}
    log("p' solves:", psolve)

    def ratio(a_, b_):
        return round(psolve[a_]["ms"] / psolve[b_]["ms"], 2) if psolve[b_]["ms"] > 0 else None

    speedup = {"learned_vs_sgs_fp64": ratio("sgs_fp64", "learned_fcg"),
               "learned_vs_plain_fp64": ratio("plain_fp64", "learned_fcg"),
               "mg_vs_sgs_fp64": ratio("sgs_fp64", "mg"), "learned_vs_mg": ratio("mg", "learned_fcg")}
    # momentum solve (3 components): plain FP64 vs FP32-inner / FP64-refinement, same system
    msolve = {}
    s.save_u0()
    s.momentum_assemble(case.time_step)
    for name, md in (("plain_fp64", L.SOLVE_PLAIN_FP64), ("ir_fp32", L.SOLVE_IR_FP32)):
        s.set_solve_opts("momentum", mode=md, rtol=1e-8, maxit=5000)
        r = s.momentum_solve()
        msolve[name] = {"ms": round(r["ms"], 3), "iters": r["iters"], "refinements": r["refinements"],
                        "relres": r["relres"], "fell_back": r["fell_back"]}
    log("momentum solves:", msolve)

    # roofline of the dominant kernel, timed live inside the timed steps (probe above)
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(f"{kname}_{a.n}")
        except Exception:
            traffic = None

    out = {
        "metric": "PISO steps/s at 128^3; p-solve ms to 1e-8 resid, learned+FP32 vs plain FP64",
        "value": round(1000.0 / ms_per_step, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"plain": "f64", "ir": "f64+f32", "learned": "f64+f32+bf16", "mg": "f64+f32", "sgs": "f64"}[a.mode],
        "data": "synthetic: lid-driven cavity from rest (no external data)",
        "config": {"workload": f"3D lid-driven cavity {a.n}^3 uniform hex, Re 1000 (mu 1e-3, lid 1), PISO 2 correctors, "
                               f"dt {case.time_step:g} (CFL 0.5), p' solve to true relres 1e-8 ({a.mode})",
                   "momentum": a.mom_mode, "cells": case.n_cells, "p_solve_ms": psolve,
                   "p_solve_speedup": speedup, "momentum_solve_ms": msolve,
                   "timed_steps": f"steps {a.warmup + 1}-{a.warmup + a.steps} of a run from rest",
                   "timed_p_fell_back_steps": fell_steps, "timed_not_converged_steps": not_conv_steps},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "ms_per_launch": round(kms, 5), "bytes_per_launch": kbytes, "launches_timed": klaunches},
    }
    if not a.no_full_run:
        log("the configured run: 100 steps from rest ...")
        out["config"]["full_run_100_steps"] = full_run(a.n, mode, inner[a.mode], net, mom_mode, a.fcg_post)
    if not a.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(a.n, steps=4, warmup=a.warmup)
        log("field errors against the oracle ...")
        out["field_error_vs_oracle"] = field_errors(a.n, mode, inner[a.mode], net, mom_mode, post=a.fcg_post)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
