"""The multigrid-preconditioned BiCGStab of the p' systems (CAFFE_SOLVE_MG) on the GPU.

The cycle changes how a p' system is solved, not which system is solved.  Each case runs
the solver with its p' solves in the multigrid mode, to a true FP64 relative residual of
1e-12, against the FP64 oracle: the same bars as the plain solves (fields within 1e-6).
Each case also checks that the multigrid path ran (no fallback to Jacobi BiCGStab).

Cases:
  * the immersed-sphere channel: nonsymmetric ghost rows, a fixed-pressure outlet;
  * the closed cavity: a singular pure-Neumann system, so the coarsest level is grounded;
  * the closed cavity with an immersed sphere: ghost rows, and coarse levels singular;
  * a cavity of odd sizes 13 x 7 x 5: uneven aggregates, and directions of 2 cells or
    fewer are not coarsened;
  * the skewed mesh: general metrics.
On the channel's p' system the multigrid iteration count is also compared with Jacobi
BiCGStab's.
"""
import numpy as np
import pytest

from caffe_old_amd import _lib as L
from caffe_old_amd import case as cs
from tests.helpers import interior, rel_l2

pytestmark = pytest.mark.gpu


def _pair(case, uniform=None):
    from caffe_old_amd.solver import make_solver
    from oracle.pyoracle import Oracle
    g = make_solver(case, uniform=uniform)
    g.set_solve_opts("momentum", rtol=1e-12, maxit=20000)
    g.set_solve_opts("pcorr", mode=L.SOLVE_MG, rtol=1e-12, maxit=200000)
    o = Oracle(case, rtol=1e-13, n_threads=4, n_blocks=4)
    return g, o


def _run(case, uniform=None):
    g, o = _pair(case, uniform)
    for _ in range(case.n_solves):
        g.solve()
        o.solve()
        p = g.reports()[1]
        assert p["fell_back"] == 0 and p["converged"] == 1, p
    sh = g.shape
    ug, uo = interior(g.get("u"), sh), interior(o.get("u"), sh)
    pg, po = interior(g.get("p"), sh), interior(o.get("p"), sh)
    errs = {c: rel_l2(ug[i], uo[i]) for i, c in enumerate("uvw") if np.linalg.norm(uo[i]) > 0}
    errs["p"] = rel_l2(pg - pg.mean(), po - po.mean())
    return g, o, errs


def test_mg_ib_channel():
    case = cs.channel_ib_sphere(64, 16, 16, steps=2)
    case.sphere = (0.2, (1.0, 0.5, 0.5))
    print(f"case: IB channel 64x16x16, multigrid p' solves", flush=True)
    g, o, errs = _run(case, uniform=True)
    assert max(errs.values()) <= 1e-6, errs


def test_mg_closed_cavity_singular():
    case = cs.lid_cavity(10, 0.01, 1.0, "simple", alpha_u=0.7, alpha_p=0.3, max_iterations=6)
    print("case: closed cavity 10^3 (singular p'), multigrid p' solves", flush=True)
    g, o, errs = _run(case, uniform=True)
    assert max(errs.values()) <= 1e-6, errs


def test_mg_ib_closed_cavity():
    """Ghost rows in a closed box: no face fixes the pressure, so the coarse levels are
    singular and grounded, while the ghost rows that reach into the solid anchor the fine
    system."""
    case = cs.lid_cavity(16, 1e-2, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.02, n_pcorr=2,
                         max_iterations=3)
    case.algo = "ibpiso"
    case.sphere = (0.2, (0.5, 0.45, 0.5))
    print("case: closed cavity 16^3 with an immersed sphere, multigrid p' solves", flush=True)
    g, o, errs = _run(case, uniform=True)
    p = g.reports()[1]
    assert p["iters"] <= 30, p
    assert max(errs.values()) <= 1e-6, errs


def test_mg_odd_sizes():
    nodes = cs.cartesian_nodes(1.0, 1.0, 1.0, 13, 7, 5)
    case = cs.lid_cavity(13, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.02, n_pcorr=2,
                         max_iterations=4, nodes=nodes)
    print("case: cavity 13x7x5, multigrid p' solves", flush=True)
    g, o, errs = _run(case, uniform=True)
    assert max(errs.values()) <= 1e-6, errs


def test_mg_skewed_mesh():
    nodes = cs.perturb_interior(cs.cartesian_nodes(1, 1, 1, 12, 12, 12), 0.15, 1 / 12, seed=0)
    case = cs.lid_cavity(12, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.04, n_pcorr=2,
                         max_iterations=3, nodes=nodes)
    print("case: skewed cavity 12^3, multigrid p' solves", flush=True)
    g, o, errs = _run(case, uniform=False)
    assert max(errs.values()) <= 1e-6, errs


def test_mg_iterations_vs_jacobi():
    """The same ghost-row p' system solved to 1e-8 by Jacobi BiCGStab (plain FP64) and by the
    multigrid BiCGStab: both converge, and the cycle cuts the iteration count at least 5x."""
    from caffe_old_amd.solver import make_solver
    case = cs.channel_ib_sphere(128, 32, 32, steps=3)
    print("case: IB channel 128x32x32, p' system after 2 steps", flush=True)
    g = make_solver(case)
    g.set_solve_opts("momentum", rtol=1e-8, maxit=5000)
    g.set_solve_opts("pcorr", mode=L.SOLVE_MG, rtol=1e-8, maxit=50000)
    for _ in range(2):
        g.solve()
    res = {}
    for name, md in (("plain", L.SOLVE_PLAIN_FP64), ("mg", L.SOLVE_MG)):
        g.pcorr_assemble()
        res[name] = g.pcorr_solve(mode=md, rtol=1e-8, maxit=50000)
    print(res, flush=True)
    assert res["plain"]["converged"] == 1 and res["mg"]["converged"] == 1, res
    assert res["mg"]["fell_back"] == 0
    assert 5 * res["mg"]["iters"] <= res["plain"]["iters"], res


@pytest.mark.parametrize("name", ["ib_channel", "cavity"])
def test_mg_graph_replays_match_eager(name):
    """The multigrid BiCGStab replays its iterations as a captured HIP graph; with a kernel
    probe recording, the same iterations are launched one by one.  The same p' system is
    solved several times, alternating the two, with the eager solves and synchronisations
    in between: every graph replay gives the eager solution bit for bit."""
    from caffe_old_amd.solver import make_solver
    if name == "ib_channel":
        case = cs.channel_ib_sphere(64, 16, 16, steps=2)
        case.sphere = (0.2, (1.0, 0.5, 0.5))
    else:
        case = cs.lid_cavity(24, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.02, n_pcorr=2,
                             max_iterations=2)
    print(f"case: {name}, graph replays against eager launches", flush=True)
    g = make_solver(case)
    g.set_solve_opts("momentum", rtol=1e-10, maxit=5000)
    g.set_solve_opts("pcorr", mode=L.SOLVE_PLAIN_FP64, rtol=1e-10, maxit=50000)
    g.solve()
    ref = None
    for k in range(6):
        eager = k % 2 == 1
        g.kernel_probe("mg_post0", eager)
        g.pcorr_assemble()
        r = g.pcorr_solve(mode=L.SOLVE_MG, rtol=1e-10, maxit=50000)
        g.sync()
        x = g.get("pCorr").copy()
        assert r["converged"] == 1 and r["fell_back"] == 0, r
        if ref is None:
            ref = (x, r["iters"])
        assert r["iters"] == ref[1] and np.array_equal(x, ref[0]), (k, r)
    g.kernel_probe("mg_post0", False)
    g.close()
