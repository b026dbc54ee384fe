"""The oracle's Diffusion solver (Diffusion.cpp:33-131) against closed forms.

Steady diffusion between two fixed faces with zero-gradient / empty faces elsewhere is
linear in the coordinate normal to the fixed faces.  The discrete system (interior d = A/h,
boundary d = A/(h/2)) is exact for a linear field, so the cell values are the line at the
cell centres and the extrapolated-face gradient is its slope.  Unsteady mode keeps the
reference's sign (a0 = -sum a_nb + V/dt, a_nb = +d: a step of backward diffusion), checked
on the assembled matrix."""
import numpy as np
import pytest

from caffe_old_amd import case as cs
from oracle.pyoracle import Oracle, interior


def _diff_case(n=(8, 5, 4), west=1.0, east=0.0, others="zeroGradient", steady=True, dt=0.1, phi0=0.0, nodes=None):
    ni, nj, nk = n
    b = [cs.Boundary("fixed", (0, 0, 0), east), cs.Boundary("fixed", (0, 0, 0), west)]
    b += [cs.Boundary(others, (0, 0, 0), 0.0) for _ in range(4)]
    if nodes is None:
        nodes = cs.cartesian_nodes(2.0, 1.0, 0.5, ni, nj, nk)
    return cs.Case(nodes=nodes, algo="diffusion", time_accurate=not steady, boundaries=b, p_init=phi0,
                   time_step=dt, max_iterations=2)


@pytest.mark.parametrize("others", ["zeroGradient", "empty"])
def test_steady_linear_profile(others):
    c = _diff_case(others=others)
    o = Oracle(c, rtol=1e-14)
    res = o.solve()
    ni, nj, nk = c.shape
    phi = interior(o.get("phi"), c.shape)
    xc = (np.arange(ni) + 0.5) * (2.0 / ni)
    expect = 1.0 - xc / 2.0
    assert np.allclose(phi, expect[None, None, :], rtol=0, atol=1e-12)
    g = interior(o.get("gradPhi"), c.shape)
    assert np.allclose(g[0], -0.5, atol=1e-11) and np.allclose(g[1], 0.0, atol=1e-11) and np.allclose(g[2], 0.0, atol=1e-11)
    assert res < 1e-12


def test_matrix_rows_and_unsteady_sign():
    c = _diff_case(n=(4, 3, 3), steady=False, dt=0.25, phi0=2.0)
    o = Oracle(c, rtol=1e-14)
    o.solve()
    rp, cols, vals = o.matrix(0)
    n = len(rp) - 1
    A = np.zeros((n, n))
    for r in range(n):
        A[r, cols[rp[r]:rp[r + 1]]] = vals[rp[r]:rp[r + 1]]
    b = o.rhs(0)
    assert np.allclose(A, A.T, rtol=0, atol=0)          # dE(i) == dW(i+1) bitwise
    hx, hy, hz = 0.5, 1.0 / 3, 0.5 / 3
    vol = hx * hy * hz
    a0P = vol / 0.25
    # an interior cell (1,1,1): six neighbours, a0 = -(sum d) + V/dt, b = V/dt * phi_old
    r = 1 + 1 * 4 + 1 * 12
    d = [hy * hz / hx] * 2 + [hx * hz / hy] * 2 + [hx * hy / hz] * 2
    assert A[r, r] == pytest.approx(-sum(d) + a0P, rel=1e-14)
    assert A[r, r + 1] == pytest.approx(d[0], rel=1e-14) and A[r, r - 4] == pytest.approx(d[2], rel=1e-14)
    assert b[r] == pytest.approx(a0P * 2.0, rel=1e-14)
    # west boundary cell (0,1,1), fixed phi = 1: the face coefficient stays in a0 and moves
    # phi_b into b (PrimitiveBoundaryCondition.tpp:144-170)
    r0 = 0 + 4 + 12
    assert A[r0, r0] == pytest.approx(-(2 * d[0] + d[0] + 2 * d[2] + 2 * d[4]) + a0P, rel=1e-14)
    assert b[r0] == pytest.approx(a0P * 2.0 - 2 * d[0] * 1.0, rel=1e-14)
