"""The red-black symmetric Gauss-Seidel preconditioned CG (CAFFE_SOLVE_SGS_FP64), the plain
FP64 judge as strong as the reference's PCSOR(1, 1) (SparseMatrix.cpp:97-98):
  * PISO steps of a 24^3 cavity with SGS p' solves reproduce the oracle's fields within 1e-6;
  * on the same p' system it reaches the tolerance in fewer iterations than Jacobi CG, and
    both solutions agree within 1e-7;
  * a channel with a fixed-pressure outlet (nonsingular system) agrees with the plain solve;
  * ghost-row systems go to the plain BiCGStab (fell_back = 1) and still converge."""
import numpy as np
import pytest

from caffe_old_amd import case as cs
from tests.helpers import field_errors, interior, rel_l2

pytestmark = pytest.mark.gpu


def _solver(case, pc_mode, rtol=1e-10):
    from caffe_old_amd import _lib as L
    from caffe_old_amd.solver import make_solver
    g = make_solver(case)
    g.set_solve_opts("momentum", mode=L.SOLVE_PLAIN_FP64, rtol=1e-12, maxit=20000)
    g.set_solve_opts("pcorr", mode=pc_mode, rtol=rtol, maxit=200000)
    return g


def test_sgs_piso_fields_and_iterations():
    from caffe_old_amd import _lib as L
    from oracle.pyoracle import Oracle
    case = cs.lid_cavity(24, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.5 / 24, n_pcorr=2,
                         max_iterations=4)
    g = _solver(case, L.SOLVE_SGS_FP64, 1e-11)
    o = Oracle(case, rtol=1e-13, n_threads=4, n_blocks=4)
    for _ in range(case.n_solves):
        g.solve()
        o.solve()
        assert g.reports()[1]["fell_back"] == 0
    errs = field_errors(g, o)
    assert max(errs.values()) <= 1e-6, errs
    # one p' system, both judges to 1e-8
    g.save_u0()
    g.momentum_assemble(case.time_step)
    g.momentum_solve()
    g.rhie_chow()
    g.pcorr_assemble()
    rs = g.pcorr_solve(mode=L.SOLVE_SGS_FP64, rtol=1e-8, maxit=100000)
    xs = interior(g.get("pCorr"), g.shape).copy()
    rj = g.pcorr_solve(mode=L.SOLVE_PLAIN_FP64, rtol=1e-8, maxit=100000)
    xj = interior(g.get("pCorr"), g.shape).copy()
    print(f"24^3 p' to 1e-8: SGS-PCG {rs['iters']} iterations ({rs['ms']:.2f} ms), Jacobi CG {rj['iters']} ({rj['ms']:.2f} ms)")
    assert rs["converged"] and rj["converged"] and rs["fell_back"] == 0
    assert rs["iters"] < 0.75 * rj["iters"]
    assert rel_l2(xs, xj) <= 1e-7


def test_sgs_channel_with_outlet():
    from caffe_old_amd import _lib as L
    case = cs.channel_ib_sphere(48, 12, 12, steps=2)
    case.algo = "piso"
    case.sphere = None
    a = _solver(case, L.SOLVE_SGS_FP64)
    b = _solver(case, L.SOLVE_PLAIN_FP64)
    for _ in range(case.n_solves):
        a.solve()
        b.solve()
    assert a.reports()[1]["fell_back"] == 0
    ua, ub = interior(a.get("u"), a.shape), interior(b.get("u"), b.shape)
    assert rel_l2(ua, ub) <= 1e-7
    assert rel_l2(interior(a.get("p"), a.shape), interior(b.get("p"), b.shape)) <= 1e-7


def test_sgs_ghost_rows_fall_back():
    from caffe_old_amd import _lib as L
    case = cs.channel_ib_sphere(48, 12, 12, steps=1)
    case.sphere = (0.2, (1.0, 0.5, 0.5))
    g = _solver(case, L.SOLVE_SGS_FP64, 1e-8)
    g.solve()
    p = g.reports()[1]
    assert p["fell_back"] == 1 and p["converged"] == 1
