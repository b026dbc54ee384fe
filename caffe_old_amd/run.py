#!/usr/bin/env python3
"""Run a reference case directory unchanged on the GPU solver.

    python -m caffe_old_amd.run hexamesh CASE_DIR
    python -m caffe_old_amd.run {simple,piso,ibpiso} CASE_DIR [--pcorr plain|ir|learned|mg] [--momentum plain|ir]
    python -m caffe_old_amd.run diffusion CASE_DIR
    python -m caffe_old_amd.run coupled CASE_DIR [--pcorr plain|mg]

``hexamesh`` does what Modules/hexaMesh.cpp:8-28 does: read mesh/structuredMesh.info,
generate the nodes, check the faces are planar, and write mesh/structuredMesh.dat.
The solver commands follow Modules/caffeSimple.cpp, caffePiso.cpp, caffeIbPiso.cpp,
caffeDiffusion.cpp and caffeCoupled.cpp: read case/case.info, case/initialConditions.info and
mesh/structuredMesh.dat, run the RunControl loop (RunControl.cpp:53-84) with one
solver call per iteration, and write solution/<mesh name>.dat every
fileWriteInterval iterations and once more at the end, with the fields the
reference's solvers register for output: p and massFlow (scalars) and u (vector), or
for the Diffusion solver phi and gradPhi (Diffusion.cpp:21-22).

The solves are converged to a true FP64 relative residual (--rtol, default 1e-8); the
reference's PETSc solves stop at rtol 1e-5 on the preconditioned norm.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from caffe_old_amd import caseio  # noqa: E402


def hexamesh(case_dir: str) -> str:
    mesh_dir = os.path.join(case_dir, "mesh")
    name, verts, res = caseio.read_mesh_info(os.path.join(mesh_dir, "structuredMesh.info"))
    nodes = caseio.generate_mesh(verts, res)
    caseio.check_mesh(verts)
    out = os.path.join(mesh_dir, "structuredMesh.dat")
    caseio.write_mesh_file(out, name, nodes)
    print(f"HexaMeshGen: wrote {out} ({res[0]} x {res[1]} x {res[2]} nodes)")
    return out


def cell_fields(solver):
    """Interior cell values of the output fields, (nK, nJ, nI) per component."""
    nI, nJ, nK = solver.shape

    def inner(a):
        return a.reshape(nK + 2, nJ + 2, nI + 2)[1:-1, 1:-1, 1:-1]

    if solver.case.algo == "diffusion":
        g = solver.get("gradPhi")
        return {"phi": inner(solver.get("phi"))}, {"gradPhi": [inner(g[c]) for c in range(3)]}
    u = solver.get("u")
    scal = {"p": inner(solver.get("p")), "massFlow": inner(solver.get("massFlow"))}
    vec = {"u": [inner(u[c]) for c in range(3)]}
    return scal, vec


def run(algo: str, case_dir: str, pcorr: str = "plain", momentum: str = "plain", rtol: float = 1e-8,
        quiet: bool = False, max_solves: int | None = None):
    """Run the case; returns (solver, RunControl, path of the solution file).  The caller
    closes the solver."""
    from caffe_old_amd import _lib as L
    from caffe_old_amd.solver import make_solver

    name, case, rc = caseio.load_case(case_dir, algo)
    s = make_solver(case)
    modes = {"plain": L.SOLVE_PLAIN_FP64, "ir": L.SOLVE_IR_FP32, "learned": L.SOLVE_LEARNED_FCG, "mg": L.SOLVE_MG}
    s.set_solve_opts("momentum", mode=modes[momentum], rtol=rtol, maxit=5000)
    s.set_solve_opts("pcorr", mode=modes[pcorr], rtol=rtol, maxit=200000, inner_rtol=1e-4 if pcorr == "learned" else 1e-2)
    if pcorr == "learned":
        from caffe_old_amd.learned import PressureNet, trained_weights_path
        nI, nJ, nK = case.shape
        w = trained_weights_path(nI)
        if nI == nJ == nK and os.path.exists(w):
            s._net = PressureNet(seed=0).load(w)
            s.attach_net(s._net)
        elif not quiet:
            print(f"no trained network for {nI}x{nJ}x{nK}: learned solves fall back to plain FP64")
    sol_dir = os.path.join(case_dir, "solution")
    os.makedirs(sol_dir, exist_ok=True)
    out_vars = (["phi"], ["gradPhi"]) if algo == "diffusion" else (["p", "massFlow"], ["u"])
    writer = caseio.Tec360Writer(os.path.join(sol_dir, name + ".dat"), name, case.nodes, *out_vars)
    t0 = time.perf_counter()
    n = 0
    while rc.continue_run():
        rc.residual_norm = s.solve(rc.time_step)
        n += 1
        if not quiet:
            s.displayUpdateMessage()
            print(rc.update_message(), flush=True)
        if rc.write_to_file():
            writer.write(rc.sim_time, *cell_fields(s))
        if max_solves is not None and n >= max_solves:
            break
    writer.write(rc.sim_time, *cell_fields(s))
    if not quiet:
        print(f"Iterations completed: {rc.itrs}\nSimulation time: {caseio._g(rc.sim_time, 6)}\n"
              f"Elapsed time: {time.perf_counter() - t0:.3f} s")
    return s, rc, writer.path


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("module", choices=["hexamesh", "simple", "piso", "ibpiso", "diffusion", "coupled"])
    ap.add_argument("case_dir", nargs="?", default=".")
    ap.add_argument("--pcorr", choices=["plain", "ir", "learned", "mg"], default="plain")
    ap.add_argument("--momentum", choices=["plain", "ir"], default="plain")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args(argv)
    if a.module == "hexamesh":
        hexamesh(a.case_dir)
        return 0
    s, rc, path = run(a.module, a.case_dir, a.pcorr, a.momentum, a.rtol, a.quiet)
    print(f"solution written to {path}")
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
