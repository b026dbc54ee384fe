"""Momentum solves as FP32 BiCGStab rounds inside FP64 refinement (CAFFE_SOLVE_IR_FP32
for the momentum system), against the plain FP64 BiCGStab and against the oracle.

Both modes converge the three momentum systems to the same true FP64 relative residual,
so the fields after PISO steps agree to the size of that tolerance.  Ghost-row (IB)
momentum systems run the FP32 rounds too, their ghost rows in FP64 values; they are held
to the same agreement with the plain solve and, on a small channel, with the oracle.
"""
import numpy as np
import pytest

from caffe_old_amd import _lib as L
from caffe_old_amd import case as cs
from tests.helpers import interior

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _piso(case, mom_mode, steps, rtol):
    from caffe_old_amd.solver import make_solver
    s = make_solver(case)
    s.set_solve_opts("momentum", mode=mom_mode, rtol=rtol, maxit=5000)
    s.set_solve_opts("pcorr", mode=L.SOLVE_PLAIN_FP64, rtol=rtol, maxit=100000)
    reps = []
    for _ in range(steps):
        s.solve()
        reps.append(s.reports()[0])
    out = {f: s.get(f).copy() for f in ("u", "p")}
    s.close()
    return out, reps


def test_momentum_ir_matches_plain_64():
    c = cs.lid_cavity(64, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.5 / 64,
                      n_pcorr=2, max_iterations=4)
    a, ra = _piso(c, L.SOLVE_IR_FP32, 3, 1e-10)
    b, rb = _piso(c, L.SOLVE_PLAIN_FP64, 3, 1e-10)
    for r in ra:
        print(f"IR momentum: iters {r['iters']} rounds {r['refinements']} relres {r['relres']:.2e} {r['ms']:.2f} ms")
        assert r["converged"] and r["relres"] <= 1e-10 and r["fell_back"] == 0 and r["refinements"] >= 2
    print("plain momentum ms", [round(r["ms"], 2) for r in rb])
    for f in ("u", "p"):
        e = _rel(a[f], b[f])
        print(f"{f}: IR vs plain rel L2 {e:.2e}")
        assert e <= 1e-8


@pytest.mark.parametrize("shape", [(128, 128, 128), (128, 64, 48)])
def test_momentum_column_kernels(shape):
    """The momentum BiCGStab's column-tile kernels on grids where a thread walks several k
    planes (128^3: 8 per thread; 128 x 64 x 48: 2, with the last column segments partly
    empty).  Both momentum modes run them (FP32 rounds and plain FP64), so each solution is
    checked on its own: the relative residual b - A x of every component, recomputed here in
    numpy from the downloaded band and right-hand side, must meet the solve's rtol; and the
    two modes' solutions must agree."""
    from caffe_old_amd.solver import make_solver
    ni, nj, nk = shape
    nodes = cs.cartesian_nodes(1.0, 1.0, 1.0, ni, nj, nk)
    c = cs.lid_cavity(ni, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.5 / ni,
                      n_pcorr=2, max_iterations=4, nodes=nodes)
    res = {}
    for mode in (L.SOLVE_IR_FP32, L.SOLVE_PLAIN_FP64):
        s = make_solver(c)
        s.set_solve_opts("momentum", mode=mode, rtol=1e-10, maxit=5000)
        s.set_solve_opts("pcorr", mode=L.SOLVE_MG, rtol=1e-10, maxit=100000)
        for _ in range(2):
            s.solve()
        s.save_u0()
        s.momentum_assemble(c.time_step)
        r = s.momentum_solve()
        A = s.matrix(0)            # (7, P): aP aE aW aN aS aT aB
        b = s.rhs(0)               # (3, P)
        x = s.get("u")             # the solution copied into u for the non-inactive cells
        sJ, sK = ni + 2, (ni + 2) * (nj + 2)
        worst = 0.
        for q in range(3):
            # the solve's iterate is zero on the halo (folded boundary columns carry no value)
            xs = np.zeros((nk + 2, nj + 2, ni + 2))
            xs[1:-1, 1:-1, 1:-1] = interior(x[q], (ni, nj, nk))
            xs = xs.ravel()
            Ax = A[0] * xs
            for m, off in ((1, 1), (2, -1), (3, sJ), (4, -sJ), (5, sK), (6, -sK)):
                Ax = Ax + A[m] * np.roll(xs, -off)
            ri = interior(b[q] - Ax, (ni, nj, nk))
            bi = interior(b[q], (ni, nj, nk))
            worst = max(worst, float(np.linalg.norm(ri) / np.linalg.norm(bi)))
        print(f"{shape} mode {mode}: iters {r['iters']} reported relres {r['relres']:.2e}, recomputed {worst:.2e}")
        assert r["converged"] and r["relres"] <= 1e-10 and worst <= 2e-10
        res[mode] = interior(x, (ni, nj, nk)).copy()
        s.close()
    e = _rel(res[L.SOLVE_IR_FP32], res[L.SOLVE_PLAIN_FP64])
    print(f"{shape}: IR vs plain momentum solution rel L2 {e:.2e}")
    assert e <= 1e-8


def test_momentum_ir_matches_oracle_12():
    from oracle.pyoracle import Oracle
    from caffe_old_amd.solver import make_solver
    c = cs.lid_cavity(12, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.5 / 12,
                      n_pcorr=2, max_iterations=3)
    g = make_solver(c)
    g.set_solve_opts("momentum", mode=L.SOLVE_IR_FP32, rtol=1e-12, maxit=5000)
    g.set_solve_opts("pcorr", rtol=1e-12, maxit=50000)
    o = Oracle(c, rtol=1e-13)
    for _ in range(c.n_solves):
        g.solve()
        o.solve()
    for f in ("u", "p"):
        e = _rel(interior(g.get(f), g.shape), interior(o.get(f), g.shape))
        print(f"{f}: IR momentum vs oracle rel L2 {e:.2e}")
        assert e <= 1e-6
    g.close()


def test_momentum_ir_ghost_rows():
    case = cs.channel_ib_sphere(64, 16, 16, steps=3)
    case.sphere = (0.2, (1.0, 0.5, 0.5))
    a, ra = _piso(case, L.SOLVE_IR_FP32, 3, 1e-10)
    b, rb = _piso(case, L.SOLVE_PLAIN_FP64, 3, 1e-10)
    for r in ra:
        print(f"IR momentum, ghost rows: iters {r['iters']} rounds {r['refinements']} relres {r['relres']:.2e} "
              f"fell_back {r['fell_back']} {r['ms']:.2f} ms")
        assert r["converged"] and r["relres"] <= 1e-10
    assert any(r["fell_back"] == 0 for r in ra)   # the FP32 rounds carried at least one solve
    for f in ("u", "p"):
        e = _rel(a[f], b[f])
        print(f"{f}: IR vs plain rel L2 {e:.2e}")
        assert e <= 1e-8


def test_momentum_ir_ghost_rows_match_oracle():
    from oracle.pyoracle import Oracle
    from caffe_old_amd.solver import make_solver
    case = cs.channel_ib_sphere(64, 16, 16, steps=2)
    case.sphere = (0.2, (1.0, 0.5, 0.5))
    g = make_solver(case)
    g.set_solve_opts("momentum", mode=L.SOLVE_IR_FP32, rtol=1e-12, maxit=5000)
    g.set_solve_opts("pcorr", rtol=1e-12, maxit=200000)
    o = Oracle(case, rtol=1e-13, n_threads=4, n_blocks=4)
    for _ in range(case.n_solves):
        g.solve()
        o.solve()
    for f in ("u", "p"):
        gv, ov = interior(g.get(f), g.shape), interior(o.get(f), g.shape)
        if f == "p":
            gv, ov = gv - gv.mean(), ov - ov.mean()
        e = _rel(gv, ov)
        print(f"{f}: IR momentum with ghost rows vs oracle rel L2 {e:.2e}")
        assert e <= 1e-6
    g.close()
