"""The Coupled solver with the block-triangular multigrid preconditioner (CAFFE_SOLVE_MG) on
the GPU.  The preconditioner changes how the 4-variable system of Coupled.cpp:54-234 is
solved, not which system: each case runs the solver in that mode to a true FP64 relative
residual of 1e-13 against the FP64 oracle, to the bars of the Jacobi-preconditioned solve
(fields within 1e-9), and checks that the multigrid path ran (no fall-back).

Cases: the 12^3 closed cavity (singular continuity rows: the coarsest level is grounded),
the unsteady 16 x 16 x 1 cavity (one cell deep: that direction is never coarsened), the
skewed 4^3 cavity (non-planar faces), the uniform channel with a fixed-pressure outlet
(its closed form), and the iteration count against Jacobi BiCGStab on the 32^3 cavity."""
import numpy as np
import pytest

from caffe_old_amd import _lib as L
from caffe_old_amd import case as cs
from tests.helpers import interior
from tests.test_gpu_coupled import _compare_fields, _skewed_nodes
from tests.test_oracle_coupled import channel

pytestmark = pytest.mark.gpu


def _pair(case, uniform=None):
    from caffe_old_amd.solver import make_solver
    from oracle.pyoracle import Oracle
    g = make_solver(case, uniform=uniform)
    g.set_solve_opts("pcorr", mode=L.SOLVE_MG, rtol=1e-13, maxit=200000)
    return g, Oracle(case, rtol=1e-14)


@pytest.mark.parametrize("which", ["cavity3d", "cavity2d", "skewed"])
def test_mg_fields_against_oracle(which):
    if which == "cavity2d":
        b = cs._walls()
        b[4], b[5] = cs.Boundary("empty"), cs.Boundary("empty")
        c = cs.Case(nodes=cs.cartesian_nodes(1.0, 1.0, 0.1, 16, 16, 1), algo="coupled", time_accurate=True, mu=0.01,
                    alpha_u=0.7, n_inner=2, boundaries=b, time_step=0.05)
        uniform, iters = True, 6
    elif which == "cavity3d":
        c = cs.lid_cavity(12, 0.01, algo="coupled")
        uniform, iters = True, 5
    else:
        c = cs.lid_cavity(4, 0.01, algo="coupled", nodes=_skewed_nodes())
        uniform, iters = False, 4
    print(f"case: Coupled {which}, multigrid block preconditioner", flush=True)
    g, o = _pair(c, uniform=uniform)
    for it in range(iters):
        eg = g.solve()
        eo = o.solve()
        p = g.reports()[1]
        assert p["fell_back"] == 0 and p["converged"] == 1, (it, p)
        assert eg == pytest.approx(eo, rel=1e-8), (it, eg, eo)
    print(f"{which}: last solve {p['iters']} iterations, relres {p['relres']:.1e}, {p['ms']:.2f} ms", flush=True)
    _compare_fields(g, o, c.shape, 1e-9)
    g.close()


def test_mg_uniform_channel_flow():
    c = channel(om=1.0)
    print("case: Coupled uniform channel, multigrid block preconditioner", flush=True)
    g, _ = _pair(c, uniform=True)
    for _ in range(3):
        e = g.solve()
        assert g.reports()[1]["fell_back"] == 0
    u, p = interior(g.get("u"), c.shape), interior(g.get("p"), c.shape)
    assert np.abs(u[0] - 1.0).max() < 1e-11 and np.abs(u[1:]).max() < 1e-11 and np.abs(p).max() < 1e-11
    assert e < 1e-10
    g.close()


def test_mg_iterations_vs_jacobi():
    """The same Coupled system (32^3 cavity, second iteration) solved to 1e-8 by Jacobi
    BiCGStab and by the block-triangular multigrid BiCGStab: both converge, and the
    multigrid one takes at least 4x fewer iterations."""
    from caffe_old_amd.solver import make_solver
    res = {}
    for name, md in (("plain", L.SOLVE_PLAIN_FP64), ("mg", L.SOLVE_MG)):
        c = cs.lid_cavity(32, 0.01, algo="coupled")
        g = make_solver(c)
        g.set_solve_opts("pcorr", mode=md, rtol=1e-8, maxit=200000)
        print(f"case: Coupled 32^3 cavity, {name}", flush=True)
        for _ in range(2):
            g.solve()
        res[name] = g.reports()[1]
        g.close()
    print(res, flush=True)
    assert res["plain"]["converged"] == 1 and res["mg"]["converged"] == 1, res
    assert res["mg"]["fell_back"] == 0
    assert 4 * res["mg"]["iters"] <= res["plain"]["iters"], res


def test_mg_grid_beyond_one_round_of_blocks():
    """80^3 cavity (512000 cells): the multigrid BiCGStab kernels launch 500 blocks, more than
    the chip holds at once, so late blocks start after block 0 has updated the solver state.
    Two Coupled iterations in the multigrid mode reproduce the Jacobi-preconditioned solve's
    fields within 1e-9, converged and without falling back."""
    from caffe_old_amd.solver import make_solver
    assert 80 ** 3 // 1024 > 256
    fields = {}
    for name, md in (("plain", L.SOLVE_PLAIN_FP64), ("mg", L.SOLVE_MG)):
        c = cs.lid_cavity(80, 0.01, algo="coupled")
        g = make_solver(c)
        g.set_solve_opts("pcorr", mode=md, rtol=1e-12, maxit=200000)
        print(f"case: Coupled 80^3 cavity, {name}", flush=True)
        for _ in range(2):
            g.solve()
            p = g.reports()[1]
            assert p["converged"] == 1, (name, p)
            if name == "mg":
                assert p["fell_back"] == 0, p
        fields[name] = (interior(g.get("u"), c.shape).copy(), interior(g.get("p"), c.shape).copy())
        g.close()
    (ua, pa), (ub, pb) = fields["mg"], fields["plain"]
    assert np.linalg.norm(ua - ub) <= 1e-9 * np.linalg.norm(ub)
    assert np.linalg.norm((pa - pa.mean()) - (pb - pb.mean())) <= 1e-9 * np.linalg.norm(pb - pb.mean())
