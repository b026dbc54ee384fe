"""The two other reference configurations at their full sizes on the GPU.

At these sizes the oracle would take minutes per step, so the checks are exact where the
oracle is cheap (the immersed-boundary cell statuses) and otherwise size-independent
properties of the result:
  * fields finite; SOLID interior cells at u = p = 0;
  * per-cell mass balance of the ACTIVE cells after the correctors (the p' solve to a
    true relative residual of 1e-8 leaves each cell's net outflow far below the flux
    scale rho U h^2);
  * on the perturbed mesh, the domain volume, and sampled face area vectors and cell
    volumes bit-identical to the reference's own compiled geometry code (oracle/_ref).
Each test prints its steps/s.
"""
import time

import numpy as np
import pytest

from caffe_old_amd import _lib as L
from caffe_old_amd import case as cs
from tests.helpers import interior

pytestmark = pytest.mark.gpu


def test_ib_channel_full_size():
    """Channel 4 x 1 x 1 on 256 x 64 x 64, sphere D = 0.2 at (1, .5, .5), Re_D 200, PISO."""
    from caffe_old_amd.solver import make_solver
    from oracle.pyoracle import Oracle
    case = cs.channel_ib_sphere(256, 64, 64, steps=4)
    g = make_solver(case)
    g.set_solve_opts("momentum", rtol=1e-8, maxit=5000)
    g.set_solve_opts("pcorr", mode=L.SOLVE_IR_FP32, rtol=1e-8, maxit=50000, inner_rtol=1e-2)
    g.solve()                              # first step also flags cells and builds ghost rows
    g.sync()
    t0 = time.perf_counter()
    for _ in range(2):
        g.solve()
    g.sync()
    dt = (time.perf_counter() - t0) / 2
    sh = g.shape
    st_g = interior(g.get("status"), sh).astype(int)
    o = Oracle(case)
    o.compute_ib_field()
    st_o = o.imap()[1].reshape(sh[2], sh[1], sh[0])
    assert np.array_equal(st_g, st_o)
    n_ghost = int((st_g == 1).sum())
    n_solid = int((st_g == 2).sum())
    print(f"IB channel 256x64x64: {1.0 / dt:.2f} steps/s ({dt * 1e3:.0f} ms/step), {n_ghost} ghost and {n_solid} inactive cells")
    assert n_ghost > 100 and n_solid > 100
    u = interior(g.get("u")[0], sh)
    p = interior(g.get("p"), sh)
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(p))
    assert np.all(u[st_g == 2] == 0.0) and np.all(p[st_g == 2] == 0.0)
    h = 1.0 / 64
    m = interior(g.get("massFlow"), sh)
    worst = float(np.abs(m[st_g == 0]).max()) / (1.0 * h * h)
    print(f"max ACTIVE-cell net outflow / (rho U h^2) = {worst:.2e}")
    assert worst <= 1e-6


def test_skewed_cavity_256_full_size():
    """256^3 cavity, interior nodes moved by +/-0.15 h (seeded Philox), Re 1000, PISO."""
    from caffe_old_amd.solver import make_solver
    case = cs.skewed_cavity(256, steps=3)
    g = make_solver(case, uniform=False)
    n = 256
    sh = g.shape
    vol = interior(g.metric("vol"), sh)
    assert abs(vol.sum() - 1.0) <= 1e-12
    # sampled faces and cells against the reference's own compiled geometry code
    # (oracle/_ref, built from the reference's Geometry.cpp): face area vectors as
    # HexaFvmMesh.cpp:725-729 composes them, cell volumes (Geometry.cpp:40-59)
    from oracle import pyoracle as po
    ref = po.ref_lib()
    if ref is not None:
        import ctypes as C
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        nodes = case.nodes
        P = g.P
        sJ, sK = n + 2, (n + 2) ** 2
        rng = np.random.default_rng(5)
        for nm, quad in (("SI", lambda i, j, k: (nodes[k, j, i], nodes[k, j + 1, i], nodes[k + 1, j + 1, i], nodes[k + 1, j, i])),
                         ("SJ", lambda i, j, k: (nodes[k, j, i], nodes[k + 1, j, i], nodes[k + 1, j, i + 1], nodes[k, j, i + 1])),
                         ("SK", lambda i, j, k: (nodes[k, j, i], nodes[k, j, i + 1], nodes[k, j + 1, i + 1], nodes[k, j + 1, i]))):
            S = g.metric(nm)
            for _ in range(400):
                i, j, k = (int(v) for v in rng.integers(0, n, 3))
                q = np.ascontiguousarray(np.array(quad(i, j, k)))
                out = np.zeros(3)
                ref.ref_face_area_normal(dp(q), dp(out))
                p = (i + 1) + (j + 1) * sJ + (k + 1) * sK
                assert np.array_equal(S[:, p], out), (nm, i, j, k)
            del S
        volp = g.metric("vol")
        for _ in range(400):
            i, j, k = (int(v) for v in rng.integers(0, n, 3))
            h8 = np.ascontiguousarray(np.array([nodes[k, j, i], nodes[k, j, i + 1], nodes[k, j + 1, i + 1], nodes[k, j + 1, i],
                                                nodes[k + 1, j, i], nodes[k + 1, j, i + 1], nodes[k + 1, j + 1, i + 1],
                                                nodes[k + 1, j + 1, i]]))
            assert volp[(i + 1) + (j + 1) * sJ + (k + 1) * sK] == ref.ref_hex_volume(dp(h8))
        print("skewed 256^3: 1200 sampled face area vectors and 400 cell volumes bit-identical to the reference's code")
    h2 = (1.0 / n) ** 2
    g.set_solve_opts("momentum", rtol=1e-8, maxit=5000)
    g.set_solve_opts("pcorr", mode=L.SOLVE_IR_FP32, rtol=1e-8, maxit=100000, inner_rtol=1e-2)
    g.solve()
    g.sync()
    t0 = time.perf_counter()
    g.solve()
    g.sync()
    dt = time.perf_counter() - t0
    m, pr = g.reports()
    print(f"skewed 256^3: {1.0 / dt:.2f} steps/s ({dt * 1e3:.0f} ms/step); last p' solve {pr['iters']} iterations, relres {pr['relres']:.1e}")
    assert pr["converged"]
    u = interior(g.get("u")[0], sh)
    assert np.all(np.isfinite(u))
    mf = interior(g.get("massFlow"), sh)
    worst = float(np.abs(mf).max()) / (1.0 * h2)
    print(f"max net outflow / (rho U h^2) = {worst:.2e}")
    assert worst <= 1e-6
