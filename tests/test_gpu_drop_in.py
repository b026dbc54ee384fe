"""The compiled C++ caller (tests/cpp/drop_in.cpp over include/caffe_hip.hpp) on the GPU: its
row-by-row SparseMatrix system solves to the exact solution, and its three PISO steps of the
16^3 cavity through the Solver-shaped FlowSolver match the oracle's fields within 1e-6."""
import os
import subprocess

import numpy as np
import pytest

from caffe_old_amd import case as cs
from tests.helpers import interior, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_drop_in(tmp_path):
    from oracle.pyoracle import Oracle
    exe = os.path.join(ROOT, "caffe_old_amd", "csrc", "_build", "drop_in")
    out = tmp_path / "fields.bin"
    print("running", exe, flush=True)
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "drop_in sparse ok" in r.stdout and "drop_in piso ok" in r.stdout
    n = 16
    f = np.fromfile(out, dtype=np.float64).reshape(4, n, n, n)
    case = cs.lid_cavity(n, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.5 / n, n_pcorr=2,
                         max_iterations=4)
    o = Oracle(case, rtol=1e-13)
    for _ in range(3):
        o.solve()
    uo = interior(o.get("u"), (n, n, n))
    po = interior(o.get("p"), (n, n, n))
    for q in range(3):
        if np.linalg.norm(uo[q]) > 0:
            assert rel_l2(f[q], uo[q]) <= 1e-6, q
    assert rel_l2(f[3] - f[3].mean(), po - po.mean()) <= 1e-6
