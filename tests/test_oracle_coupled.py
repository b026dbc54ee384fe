"""The oracle's Coupled solver (Coupled.cpp:38-338) against closed forms and operator
identities.

Closed forms.  Uniform channel flow (inlet u = (U, 0, 0) on the west face, a fixed-pressure
outlet on the east face, empty side faces) is an exact discrete solution: u = (U, 0, 0) and
p = 0 satisfy the upwind momentum rows, and the continuity rows balance the inflow folded
into the right-hand side.  Plane Couette flow (a wall below, a moving wall above, empty
faces elsewhere) has the linear profile u_x = U (j + 1/2) / nJ as an exact solution of the
diffusion rows (half-cell wall distance), with zero pressure.  With omega = 1 both hold
after the first iteration, and they stay exact afterwards.

Operator identities of the assembled 4-variable matrix.  Row sums: the pressure
coefficients of a momentum row add up to the sum of the face normals, which is zero for a
closed cell, with the zero-gradient fold included; the d-Laplacian rows add up to zero.  On
a mesh stretched in x the volume weights interpolate a linear field exactly at the faces,
so the momentum rows applied to p = x give V e_x (the divergence theorem) and the
continuity rows applied to u = (x, 0, 0) give rho V, in cells away from the boundary.
Swapping the weights g and 1 - g, or the sign of a normal, fails these."""
import numpy as np
import pytest
import scipy.sparse as sp

from caffe_old_amd import case as cs
from oracle.pyoracle import Oracle, interior


def _nodes(xs, ys, zs):
    z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([x, y, z], axis=-1)


def _coupled(nodes, bounds, om=1.0, rho=1.0, mu=0.05):
    return cs.Case(nodes=nodes, algo="coupled", rho=rho, mu=mu, alpha_u=om, boundaries=bounds, max_iterations=2)


def _csr(o):
    rp, cols, vals = o.matrix(2)
    n = len(rp) - 1
    return sp.csr_matrix((vals, cols, rp), shape=(n, n))


def channel(ni=8, nj=4, nk=3, U=1.0, om=1.0):
    b = [cs.Boundary("outlet", (0, 0, 0), 0.0), cs.Boundary("inlet", (U, 0, 0))] + [cs.Boundary("empty") for _ in range(4)]
    return _coupled(cs.cartesian_nodes(2.0, 1.0, 0.7, ni, nj, nk), b, om=om)


def couette(ni=3, nj=10, nk=2, U=1.5):
    b = [cs.Boundary("empty"), cs.Boundary("empty"), cs.Boundary("wall", (U, 0, 0)), cs.Boundary("wall", (0, 0, 0)),
         cs.Boundary("empty"), cs.Boundary("empty")]
    return _coupled(cs.cartesian_nodes(1.0, 1.0, 1.0, ni, nj, nk), b, mu=0.2)


def exact_channel_state(o, c, U=1.0):
    """u = (U, 0, 0) on cells and patches, I-face mass flows U * area, p = 0."""
    ni, nj, nk = c.shape
    P = (ni + 2) * (nj + 2) * (nk + 2)
    u = np.zeros((3, P))
    u[0] = U
    o.set("u", u)
    area = (1.0 / nj) * (0.7 / nk)
    o.set("mI", np.full(P, U * area))
    return area


@pytest.mark.parametrize("iters", [1, 3])
def test_uniform_channel_flow(iters):
    # omega = 1: exact from rest after the first iteration.  (This solver's pressure gradient
    # sees zero on interior p faces, so a nonzero p would drive gradients in the boundary
    # cells; the channel iteration amplifies rounding about five-fold per pass, which the
    # tolerance allows for three passes.)
    c = channel(om=1.0)
    o = Oracle(c, rtol=1e-14)
    for _ in range(iters):
        e = o.solve()
    u = interior(o.get("u"), c.shape)
    p = interior(o.get("p"), c.shape)
    assert np.abs(u[0] - 1.0).max() < 1e-11 and np.abs(u[1]).max() < 1e-11 and np.abs(u[2]).max() < 1e-11
    assert np.abs(p).max() < 1e-11
    area = (1.0 / 4) * (0.7 / 3)
    mI = interior(o.get("mI"), c.shape)
    assert np.allclose(mI, area, rtol=1e-11, atol=0)
    assert e < 1e-10


def test_uniform_flow_is_a_fixed_point_with_relaxation():
    # omega < 1: the relaxed rows (a_P / omega on the left, (1 - omega) a_P u* on the right)
    # keep the exact state
    c = channel(om=0.7)
    o = Oracle(c, rtol=1e-14)
    area = exact_channel_state(o, c)
    e = o.solve()
    u = interior(o.get("u"), c.shape)
    assert np.abs(u[0] - 1.0).max() < 1e-12 and np.abs(u[1:]).max() < 1e-12
    assert np.abs(interior(o.get("p"), c.shape)).max() < 1e-12
    assert np.allclose(interior(o.get("mI"), c.shape), area, rtol=1e-12, atol=0)
    assert e < 1e-10


def test_couette_flow():
    c = couette()
    o = Oracle(c, rtol=1e-14)
    ni, nj, nk = c.shape
    for _ in range(3):
        o.solve()
        u = interior(o.get("u"), c.shape)
        expect = 1.5 * (np.arange(nj) + 0.5) / nj
        assert np.abs(u[0] - expect[None, :, None]).max() < 1e-12
        assert np.abs(u[1]).max() < 1e-12 and np.abs(u[2]).max() < 1e-12
        assert np.abs(interior(o.get("p"), c.shape)).max() < 1e-12


def _stretched_cavity(ni=7, nj=5, nk=4):
    xs = np.concatenate([[0.0], np.cumsum(1.3 ** np.arange(ni))])
    xs /= xs[-1]
    b = [cs.Boundary("wall") for _ in range(6)]
    b[2] = cs.Boundary("wall", (1.0, 0, 0))
    return _coupled(_nodes(xs, np.linspace(0, 0.8, nj + 1), np.linspace(0, 0.6, nk + 1)), b, om=0.7)


def test_operator_identities_stretched_mesh():
    c = _stretched_cavity()
    o = Oracle(c, rtol=1e-14)
    o.solve()
    o.solve()   # the second assembly has nonzero mass flows, d and grad p
    A = _csr(o)
    ni, nj, nk = c.shape
    N = ni * nj * nk
    vol = interior(o.metric("vol"), c.shape).ravel()
    xc = interior(o.metric("xc")[0], c.shape).ravel()
    # cells away from the boundary, in row order (i fastest)
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    inner = ((i > 0) & (i < ni - 1) & (j > 0) & (j < nj - 1) & (k > 0) & (k < nk - 1)).ravel()
    assert inner.sum() > 0
    scale = np.abs(A).max()

    e_p = np.zeros(4 * N); e_p[3 * N:] = 1.0
    y = A @ e_p
    assert np.abs(y).max() < 1e-14 * scale * 10          # G 1 = 0 and L 1 = 0 in every row

    U = np.zeros(4 * N); U[:N], U[N:2 * N], U[2 * N:3 * N] = 0.3, -0.7, 1.1
    y = A @ U
    assert np.abs(y[3 * N:][inner]).max() < 1e-14 * scale * 10   # D U = 0 inside

    px = np.zeros(4 * N); px[3 * N:] = xc
    y = A @ px
    assert np.allclose(y[:N][inner], vol[inner], rtol=1e-12)      # momentum x rows: V
    assert np.abs(y[N:3 * N].reshape(2, N)[:, inner]).max() < 1e-14
    ux = np.zeros(4 * N); ux[:N] = xc
    y = A @ ux
    assert np.allclose(y[3 * N:][inner], c.rho * vol[inner], rtol=1e-12)


def test_row_structure_and_diagonal():
    c = _stretched_cavity(5, 4, 4)
    o = Oracle(c, rtol=1e-14)
    o.solve()
    rp, cols, vals = o.matrix(2)
    ni, nj, nk = c.shape
    N = ni * nj * nk
    widths = np.diff(rp)
    r = 1 + ni * (1 + nj * 1)                # an interior cell
    assert widths[r] == 14 and widths[N + r] == 14 and widths[2 * N + r] == 14 and widths[3 * N + r] == 28
    assert widths[0] == 14 - 6 and widths[3 * N] == 28 - 12   # a corner: three off-domain faces per block
    A = _csr(o)
    assert (A.diagonal() > 0).all()
