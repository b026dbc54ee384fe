"""The Diffusion solver (Diffusion.cpp:33-131) on the GPU against the oracle and closed
forms: steady diffusion (the symmetric p'-type system, CG), unsteady diffusion with the
reference's sign (indefinite, BiCGStab), the extrapolated-face gradient, the returned
||A x - b||_inf, and the FlatPlateDiffusion2D deck through the file front end."""
import os
import re
import shutil

import numpy as np
import pytest

from caffe_old_amd import _lib as L
from caffe_old_amd import case as cs
from tests.helpers import interior

pytestmark = pytest.mark.gpu

CASES = os.path.join(os.path.dirname(__file__), "golden", "cases")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _case(nodes, types, refs, steady=True, dt=0.05, phi0=0.0, iters=3):
    b = [cs.Boundary(t, (0, 0, 0), r) for t, r in zip(types, refs)]
    return cs.Case(nodes=nodes, algo="diffusion", time_accurate=not steady, boundaries=b, p_init=phi0,
                   time_step=dt, max_iterations=iters)


def _pair(case, uniform=None):
    from caffe_old_amd.solver import make_solver
    from oracle.pyoracle import Oracle
    g = make_solver(case, uniform=uniform)
    g.set_solve_opts("pcorr", rtol=1e-13, maxit=100000)
    return g, Oracle(case, rtol=1e-14)


def test_linear_profile_uniform():
    c = _case(cs.cartesian_nodes(2.0, 1.0, 0.5, 16, 6, 5), ["fixed", "fixed", "zeroGradient", "zeroGradient", "empty", "empty"],
              [0.0, 1.0, 0, 0, 0, 0], iters=2)
    g, _ = _pair(c, uniform=True)
    res = g.solve()
    ni = c.shape[0]
    phi = interior(g.get("phi"), g.shape)
    xc = (np.arange(ni) + 0.5) * (2.0 / ni)
    assert np.allclose(phi, (1.0 - xc / 2.0)[None, None, :], rtol=0, atol=1e-11)
    gp = interior(g.get("gradPhi"), g.shape)
    assert np.allclose(gp[0], -0.5, atol=1e-10) and np.allclose(gp[1], 0, atol=1e-10) and np.allclose(gp[2], 0, atol=1e-10)
    assert res < 1e-12
    g.close()


@pytest.mark.parametrize("steady", [True, False])
def test_against_oracle_skewed(steady):
    nodes = cs.perturb_interior(cs.transfinite_nodes([(0, 0, 0), (1, 0, 0), (1.3, 1, 0), (0, 1, 0), (0, 0, 0.8), (1, 0, 0.8),
                                                      (1.3, 1, 0.8), (0, 1, 0.8)], 10, 8, 6), 0.1, 0.1, seed=3)
    c = _case(nodes, ["fixed", "fixed", "zeroGradient", "fixed", "zeroGradient", "empty"], [0.0, 1.0, 0.0, 0.25, 0, 0],
              steady=steady, dt=0.02, phi0=0.5, iters=4)
    g, o = _pair(c, uniform=False)
    for _ in range(c.n_solves):
        rg = g.solve()
        ro = o.solve()
        assert rg <= 1e-10 and ro <= 1e-10, (rg, ro)
    eg = _rel(interior(g.get("phi"), g.shape), interior(o.get("phi"), g.shape))
    gg, go = interior(g.get("gradPhi"), g.shape), interior(o.get("gradPhi"), g.shape)
    egr = max(_rel(gg[q], go[q]) for q in range(3))
    m, p = g.reports()
    print(f"steady={steady}: phi rel L2 {eg:.2e}, gradPhi {egr:.2e}; last solve {p['iters']} its relres {p['relres']:.1e}")
    assert eg <= 1e-10 and egr <= 1e-8
    g.close()


def _copy_case(src, dst):
    for root, _, files in os.walk(src):
        rel = os.path.relpath(root, src)
        os.makedirs(os.path.join(dst, rel), exist_ok=True)
        for f in files:
            shutil.copyfile(os.path.join(root, f), os.path.join(dst, rel, f))


def test_flat_plate_deck(tmp_path):
    from caffe_old_amd import caseio as io
    from caffe_old_amd.run import hexamesh, run
    from oracle.pyoracle import Oracle
    d = str(tmp_path / "FlatPlateDiffusion2D")
    _copy_case(os.path.join(CASES, "FlatPlateDiffusion2D"), d)
    mi = os.path.join(d, "mesh", "structuredMesh.info")
    text = open(mi).read()
    with open(mi, "w") as f:
        f.write(re.sub(r"Resolution \([^)]*\)", "Resolution (41,21,2)", text))
    hexamesh(d)
    s, rc, path = run("diffusion", d, rtol=1e-13, quiet=True)
    sh = s.shape
    s.close()
    name, case, _ = io.load_case(d, "diffusion")
    nodes, zones = io.read_tec360(path)
    assert len(zones) == case.n_solves + 1
    o = Oracle(case, rtol=1e-14)
    for _ in range(case.n_solves):
        o.solve()
    e = _rel(zones[-1]["phi"], interior(o.get("phi"), sh))
    go = interior(o.get("gradPhi"), sh)
    eg = max(_rel(zones[-1][f"gradPhi_{c}"], go[q]) for q, c in enumerate("xy"))
    print(f"{name}: file vs oracle phi {e:.2e}, gradPhi {eg:.2e}")
    assert e <= 1e-10 and eg <= 1e-8
