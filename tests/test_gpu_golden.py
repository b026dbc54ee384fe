"""The GPU (through the C ABI) against the committed oracle golden vectors
(tests/golden/oracle_fixtures.npz):
  * 4^3 and 8^3 SIMPLE cavities, iterations 1 and 5: from the recorded state before the
    iteration the momentum band and right-hand sides, and from the recorded state after
    the momentum solve the p' coefficients and right-hand side, within 1e-12 (max-norm
    relative); u and p after 1 and 5 solves from rest within 1e-9 relative L2 (p with the
    mean removed), mass fluxes within 1e-9;
  * the ghost rows of the 16^3 cavity with an immersed sphere, within 1e-10;
  * the learned network on the 8^3 grid: forward within 5e-3 of the BF16-emulating oracle
    and 3e-2 of FP64, parameter-gradient norms per layer within 5e-2."""
import os

import numpy as np
import pytest

from tests.golden.make_fixtures import OUT, ghost_case, simple_case
from tests.helpers import STATE, interior, rel_inf, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fx():
    return np.load(OUT)


def _solver(case):
    from caffe_old_amd.solver import make_solver
    g = make_solver(case)
    g.set_solve_opts("momentum", rtol=1e-12, maxit=20000)
    g.set_solve_opts("pcorr", rtol=1e-12, maxit=200000)
    return g


@pytest.mark.parametrize("n,it", [(4, 1), (4, 5), (8, 1), (8, 5)])
def test_simple_golden(fx, n, it):
    c = simple_case(n)
    key = f"simple{n}_it{it}"
    g = _solver(c)
    st = interior(np.arange(g.P), g.shape).ravel()
    act = np.zeros(g.P, dtype=bool)
    act[st] = True
    for k in STATE:
        g.set(k, fx[f"{key}_before_{k}"])
    g.momentum_assemble(c.time_step)
    assert rel_inf(g.matrix(0), fx[f"{key}_band"], act) <= 1e-12
    for q in range(3):
        assert rel_inf(g.rhs(0)[q], fx[f"{key}_mrhs"][q], act) <= 1e-12, q
    for k in STATE:
        g.set(k, fx[f"{key}_mid_{k}"])
    g.pcorr_assemble()
    sym = g.matrix(1)
    assert rel_inf(sym[0], fx[f"{key}_pcsym"][0], act) <= 1e-12
    for a in (1, 2, 3):
        assert rel_inf(sym[a], fx[f"{key}_pcsym"][a]) <= 1e-12, a
    assert rel_inf(g.rhs(1), fx[f"{key}_pcrhs"], act) <= 1e-12
    g.close()
    # fields after `it` solves from rest
    g = _solver(c)
    for _ in range(it):
        g.solve()
    sh = g.shape
    ug, uo = interior(g.get("u"), sh), interior(fx[f"{key}_after_u"], sh)
    for q in range(3):
        assert rel_l2(ug[q], uo[q]) <= 1e-9, q
    pg, po = interior(g.get("p"), sh), interior(fx[f"{key}_after_p"], sh)
    assert rel_l2(pg - pg.mean(), po - po.mean()) <= 1e-9
    for k in ("mI", "mJ", "mK"):
        assert rel_inf(g.get(k), fx[f"{key}_after_{k}"]) <= 1e-9, k


def test_ghost_rows_golden(fx):
    from caffe_old_amd.solver import make_solver
    g = make_solver(ghost_case())
    g.ib_field()
    g.ib_coeffs()
    assert np.array_equal(interior(g.get("status"), g.shape).ravel().astype(np.int8), fx["ib16_status"])
    cells, cols, vm, vp = g.ghost_rows()
    assert np.array_equal(np.sort(cells), np.sort(fx["ib16_cells"]))
    where = {int(c): q for q, c in enumerate(fx["ib16_cells"])}
    for vals, name in ((vm, "M"), (vp, "P")):
        for q, cell in enumerate(cells):
            r = where[int(cell)]
            want = {int(c): v for c, v in zip(fx[f"ib16_cols{name}"][r], fx[f"ib16_vals{name}"][r]) if c >= 0}
            got = {int(c): v for c, v in zip(cols[q], vals[q]) if c >= 0}
            assert set(got) == set(want)
            # 1e-10: the weights invert an 8x8 monomial matrix of cell centres (condition ~1e5)
            assert max(abs(got[c] - want[c]) for c in want) <= 1e-10


def test_network_golden(fx):
    from caffe_old_amd.learned import PressureNet, trained_weights_path
    from oracle import net_oracle
    net = PressureNet(seed=0).load(trained_weights_path(128))
    y = net.forward(fx["net_x"].astype(np.float32)).astype(np.float64)
    assert rel_l2(y, fx["net_y_bf16"]) <= 5e-3
    assert rel_l2(y, fx["net_y"]) <= 3e-2
    gp = net.backward(fx["net_x"].astype(np.float32), fx["net_gout"].astype(np.float32)).astype(np.float64)
    off = np.cumsum([0] + net_oracle.layer_sizes())
    norms = np.array([np.linalg.norm(gp[off[q]:off[q + 1]]) for q in range(len(off) - 1)])
    assert np.abs(norms / fx["net_gnorm"] - 1).max() <= 5e-2, norms / fx["net_gnorm"]
    net.close()
