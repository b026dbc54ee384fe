"""Closed-form pins of the oracle's SIMPLE / PISO momentum right-hand side, boundary fold
and Rhie-Chow face flux (oracle/caffe_oracle.cpp: assembleMomentum, PrimBC::fold, rhieChow).

Every expected value here is computed from the node coordinates alone (numpy), never from
the oracle's metric arrays, so a wrong sign, a dropped term or a wrong boundary distance
in the oracle fails a test:

  * momentum right-hand side on a stretched lid cavity with a frozen state (no face flux,
    u = u* = (0, s, 0), u0 = (0, 0, t), grad p = e_x):
        b = a0P u0 - V grad p + (1 - alpha) a0 u* + mu d_wall u_wall (FIXED faces)
    with a0 = (mu sum_f d_f + a0P) / alpha and a_nb = -mu d_f (Simple.cpp:117-179,
    PrimitiveBoundaryCondition.tpp:144-255);
  * p = x on the same mesh through correct() (FvScheme.tpp:12-40, FvScalarScheme.cpp:48-66)
    gives b_x = -V in the interior cells;
  * Rhie-Chow (Simple.cpp:410-449) with h = c, d = D, p = g.x and rho on a stretched mesh:
    every face flux is rho A_f (c - D g).n_f, boundary faces included;
  * uniform channel flow (inlet U on the west face, outlet east, empty elsewhere) is a fixed
    point of SIMPLE and PISO for any relaxation: u = (U, 0, 0), p = 0, m_I = U A;
  * plane Couette flow (wall south, moving wall north, empty elsewhere) on a mesh stretched
    in y: u_x linear in the cell-centroid y is exact, after one SIMPLE iteration from rest
    with alpha = 1 and as a fixed point with alpha = 0.7.

tests/test_oracle_mutants.py rebuilds the oracle with planted mistakes and checks that
these pins catch each of them."""
import numpy as np
import pytest

from caffe_old_amd import case as cs
from oracle.pyoracle import Oracle, interior
from tests.helpers import oracle_momentum_band, oracle_rhs_padded

pytestmark = pytest.mark.usefixtures("oracle_built")


def _nodes(xs, ys, zs):
    z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([x, y, z], axis=-1)


def _stretched(n, r):
    e = np.concatenate([[0.0], np.cumsum(r ** np.arange(n))])
    return e / e[-1]


class Box:
    """Closed-form finite-volume geometry of a tensor-product box mesh."""

    def __init__(self, xs, ys, zs):
        self.e = [np.asarray(xs, float), np.asarray(ys, float), np.asarray(zs, float)]
        self.w = [np.diff(a) for a in self.e]            # cell widths
        self.c = [0.5 * (a[1:] + a[:-1]) for a in self.e]  # centroids
        self.shape = tuple(len(w) for w in self.w)

    def cellwise(self):
        wx, wy, wz = self.w
        WZ, WY, WX = np.meshgrid(wz, wy, wx, indexing="ij")
        return WX, WY, WZ

    def vol(self):
        WX, WY, WZ = self.cellwise()
        return WX * WY * WZ

    def d_faces(self, axis):
        """(d_low, d_high) per cell along axis: |S|^2 / (S . r) = A / distance, the distance
        being centroid to centroid inside, centroid to face at the boundary."""
        WX, WY, WZ = self.cellwise()
        area = [WY * WZ, WX * WZ, WX * WY][axis]
        c, e = self.c[axis], self.e[axis]
        lo = np.empty_like(c)
        hi = np.empty_like(c)
        lo[0] = c[0] - e[0]
        lo[1:] = c[1:] - c[:-1]
        hi[-1] = e[-1] - c[-1]
        hi[:-1] = c[1:] - c[:-1]
        shp = [(1, 1, -1), (1, -1, 1), (-1, 1, 1)][axis]
        return area / lo.reshape(shp), area / hi.reshape(shp), area


def _lid_case(box, algo, alpha, mu, dt=1.0, unsteady=False):
    b = [cs.Boundary("wall") for _ in range(6)]
    b[2] = cs.Boundary("wall", (1.0, 0.0, 0.0))
    return cs.Case(nodes=_nodes(*box.e), algo=algo, time_accurate=unsteady, rho=1.0, mu=mu, alpha_u=alpha, alpha_p=0.3,
                   n_pcorr=2, boundaries=b, time_step=dt, max_iterations=3)


def _cells_to_padded(o, arr3, base):
    """write a (nK, nJ, nI) array into the interior of a padded array (halo kept)"""
    out = np.array(base, copy=True)
    interior(out, o.shape)[...] = arr3
    return out


@pytest.mark.parametrize("unsteady", [False, True])
def test_momentum_rhs_closed_form(unsteady):
    box = Box(_stretched(6, 1.35), _stretched(5, 0.7), _stretched(4, 1.2) * 0.8)
    alpha, mu, dt, s, t = 0.7, 0.013, 0.05, 0.37, -0.61
    o = Oracle(_lid_case(box, "piso" if unsteady else "simple", alpha, mu, dt, unsteady))
    P = o.P
    zero = np.zeros(P)
    for n in ("mI", "mJ", "mK"):
        o.set(n, zero)
    u = o.get("u")  # patches hold the wall values (FIXED): lid (1, 0, 0) on the north face
    u = np.stack([_cells_to_padded(o, np.zeros(box.shape[::-1]), u[0]),
                  _cells_to_padded(o, np.full(box.shape[::-1], s), u[1]),
                  _cells_to_padded(o, np.zeros(box.shape[::-1]), u[2])])
    o.set("u", u)
    u0 = np.zeros((3, P))
    u0[2] = t
    o.set("u0", u0)
    gp = np.zeros((3, P))
    gp[0] = 1.0
    o.set("gradP", gp)
    o.assemble_momentum(dt)

    V = box.vol()
    a0P = (1.0 * V / dt) if unsteady else np.zeros_like(V)
    dsum = np.zeros_like(V)
    d = [box.d_faces(a) for a in range(3)]
    for lo, hi, _ in d:
        dsum = dsum + lo + hi
    a0 = (mu * dsum + a0P) / alpha
    # the north boundary cells fold the lid: b -= a_N u_lid with a_N = -mu d_N(wall)
    fold_x = np.zeros_like(V)
    fold_x[:, -1, :] = mu * d[1][1][:, -1, :] * 1.0
    expect = [-V + fold_x, (1 - alpha) * a0 * s, a0P * t]
    sh = o.shape
    for c in range(3):
        b = interior(oracle_rhs_padded(o, c, P), sh)
        scale = max(np.abs(expect[c]).max(), 1e-300)
        assert np.abs(b - expect[c]).max() <= 1e-12 * scale, f"rhs component {c}"
    band, st = oracle_momentum_band(o)
    diag = interior(band[0], sh)
    assert np.abs(diag - a0).max() <= 1e-13 * a0.max()
    # off-diagonal a_E on interior faces: -mu d_E; the wall columns are dropped (-1)
    aE = interior(band[1], sh)
    assert np.abs(aE[:, :, :-1] + mu * d[0][1][:, :, :-1]).max() <= 1e-13 * a0.max()
    assert (aE[:, :, -1] == 0).all()
    aN = interior(band[3], sh)
    assert np.abs(aN[:, :-1, :] + mu * d[1][1][:, :-1, :]).max() <= 1e-13 * a0.max()
    dcell = interior(o.get("d"), sh)
    assert np.abs(dcell - V / a0).max() <= 1e-13 * (V / a0).max()


def test_pressure_gradient_source_from_linear_p():
    """p = 2.5 x through correct() (grad p by the divergence theorem with volume-weighted
    faces), then the momentum source is -V grad p = -2.5 V e_x in the interior cells."""
    box = Box(_stretched(7, 1.3), _stretched(6, 0.75), _stretched(5, 1.1))
    o = Oracle(_lid_case(box, "simple", 0.8, 0.01))
    P = o.P
    xc = np.zeros(P)
    X = np.meshgrid(box.c[2], box.c[1], box.c[0], indexing="ij")[2]
    o.set("p", _cells_to_padded(o, 2.5 * X, xc))
    o.set("pCorr", np.zeros(P))
    o.set("gradPCorr", np.zeros((3, P)))
    for n in ("mI", "mJ", "mK"):
        o.set(n, np.zeros(P))
    o.set("u", np.zeros((3, P)))  # u = 0 on cells and patches: no u* term and no lid fold
    o.correct()
    g = interior(o.get("gradP"), o.shape)[:, 1:-1, 1:-1, 1:-1]
    assert np.abs(g[0] - 2.5).max() <= 1e-12 and np.abs(g[1:]).max() <= 1e-12
    o.assemble_momentum(1.0)
    bx = interior(oracle_rhs_padded(o, 0, P), o.shape)[1:-1, 1:-1, 1:-1]
    V = box.vol()[1:-1, 1:-1, 1:-1]
    assert np.abs(bx + 2.5 * V).max() <= 1e-12 * (2.5 * V).max()


def test_rhie_chow_closed_form():
    box = Box(_stretched(6, 1.4), _stretched(5, 0.8), _stretched(4, 1.25))
    rho, D = 1.7, 0.23
    c = np.array([0.9, -0.4, 0.25])
    g = np.array([0.6, 1.1, -0.8])
    case = _lid_case(box, "simple", 1.0, 0.01)
    case.rho = rho
    o = Oracle(case)
    P = o.P
    ni, nj, nk = o.shape
    # padded coordinates: cell centroids inside, face centroids in the halo (the patches)
    def pad_axis(cen, edge):
        return np.concatenate([[edge[0]], cen, [edge[-1]]])
    px, py, pz = (pad_axis(box.c[a], box.e[a]) for a in range(3))
    Z, Y, X = np.meshgrid(pz, py, px, indexing="ij")
    p = (g[0] * X + g[1] * Y + g[2] * Z).ravel()
    o.set("p", p)
    o.set("h", np.tile(c[:, None], (1, P)))
    o.set("d", np.full(P, D))
    o.rhie_chow()
    _, _, areaI = box.d_faces(0)
    _, _, areaJ = box.d_faces(1)
    _, _, areaK = box.d_faces(2)
    for nm, area, axis in (("mI", areaI, 0), ("mJ", areaJ, 1), ("mK", areaK, 2)):
        # face array entry at padded (i, j, k) is the low face of cell (i, j, k); take the
        # faces of the domain: (n+1) along the axis, n across
        sel = [slice(1, nk + 1), slice(1, nj + 1), slice(1, ni + 1)]
        sel[2 - axis] = slice(1, [ni, nj, nk][axis] + 2)
        m = o.get(nm).reshape(nk + 2, nj + 2, ni + 2)[tuple(sel)]
        # face areas along the axis do not depend on the position along it
        a = np.take(area, [0], axis=2 - axis)
        expect = rho * a * (c[axis] - D * g[axis])
        assert np.abs(m - expect).max() <= 1e-13 * np.abs(expect).max(), nm


def _channel(algo, alpha, U=1.3, ni=8, nj=4, nk=3, unsteady=False, dt=0.05, stretched=False):
    xs = _stretched(ni, 1.15) * 2.0 if stretched else np.linspace(0, 2.0, ni + 1)
    nodes = _nodes(xs, np.linspace(0, 1.0, nj + 1), np.linspace(0, 0.7, nk + 1))
    b = [cs.Boundary("outlet", (0, 0, 0), 0.0), cs.Boundary("inlet", (U, 0, 0))] + [cs.Boundary("empty") for _ in range(4)]
    return cs.Case(nodes=nodes, algo=algo, time_accurate=unsteady, rho=1.0, mu=0.02, alpha_u=alpha, alpha_p=0.3,
                   n_pcorr=2, boundaries=b, time_step=dt, max_iterations=6)


def _impose_channel(o, c, U):
    ni, nj, nk = c.shape
    P = o.P
    u = np.zeros((3, P))
    u[0] = U
    o.set("u", u)
    o.set("p", np.zeros(P))
    area = (1.0 / nj) * (0.7 / nk)
    o.set("mI", np.full(P, U * area))
    o.set("mJ", np.zeros(P))
    o.set("mK", np.zeros(P))
    return area


@pytest.mark.parametrize("algo,alpha,unsteady,stretched", [("simple", 0.7, False, False), ("simple", 1.0, False, True),
                                                           ("piso", 1.0, True, False), ("piso", 0.7, True, True)])
def test_uniform_channel_is_a_fixed_point(algo, alpha, unsteady, stretched):
    U = 1.3
    c = _channel(algo, alpha, U, unsteady=unsteady, stretched=stretched)
    o = Oracle(c, rtol=1e-14)
    area = _impose_channel(o, c, U)
    for _ in range(4):
        e = o.solve()
        u = interior(o.get("u"), c.shape)
        assert np.abs(u[0] - U).max() <= 1e-12 * U, "u_x"
        assert np.abs(u[1:]).max() <= 1e-12 * U, "u_y, u_z"
        assert np.abs(interior(o.get("p"), c.shape)).max() <= 1e-12 * U * U
        ni, nj, nk = c.shape
        mI = o.get("mI").reshape(nk + 2, nj + 2, ni + 2)[1:-1, 1:-1, 1:ni + 2]
        assert np.abs(mI - U * area).max() <= 1e-12 * U * area, "I-face flux"
        assert np.abs(o.get("mJ")).max() <= 1e-12 * U * area and np.abs(o.get("mK")).max() <= 1e-12 * U * area
        assert e <= 1e-10


def _couette(alpha, U=1.5, ni=3, nj=9, nk=2):
    ys = _stretched(nj, 1.25)
    nodes = _nodes(np.linspace(0, 1.0, ni + 1), ys, np.linspace(0, 0.5, nk + 1))
    b = [cs.Boundary("empty"), cs.Boundary("empty"), cs.Boundary("wall", (U, 0, 0)), cs.Boundary("wall", (0, 0, 0)),
         cs.Boundary("empty"), cs.Boundary("empty")]
    c = cs.Case(nodes=nodes, algo="simple", rho=1.0, mu=0.2, alpha_u=alpha, alpha_p=0.3, boundaries=b, max_iterations=6)
    yc = 0.5 * (ys[1:] + ys[:-1])
    return c, U * yc / ys[-1]


def test_couette_exact_after_one_iteration_from_rest():
    c, expect = _couette(1.0)
    o = Oracle(c, rtol=1e-14)
    for _ in range(3):
        o.solve()
        u = interior(o.get("u"), c.shape)
        assert np.abs(u[0] - expect[None, :, None]).max() <= 1e-12 * 1.5
        assert np.abs(u[1:]).max() <= 1e-12 * 1.5
        assert np.abs(interior(o.get("p"), c.shape)).max() <= 1e-12


def test_couette_is_a_fixed_point_with_relaxation():
    c, expect = _couette(0.7)
    o = Oracle(c, rtol=1e-14)
    P = o.P
    u = o.get("u")
    u[0] = _cells_to_padded(o, np.broadcast_to(expect[None, :, None], c.shape[::-1]), u[0])
    o.set("u", u)
    ni, nj, nk = c.shape
    # I-face flux of the profile, u_x(j) times the face area dy_j dz (the empty east and
    # west faces carry it too)
    area = np.diff(c.nodes[0, :, 0, 1]) * (0.5 / nk)
    mI = np.zeros((nk + 2, nj + 2, ni + 2))
    mI[1:-1, 1:-1, 1:ni + 2] = (expect * area)[None, :, None]
    o.set("mI", mI.ravel())
    for _ in range(3):
        o.solve()
        u = interior(o.get("u"), c.shape)
        assert np.abs(u[0] - expect[None, :, None]).max() <= 1e-12 * 1.5
        assert np.abs(u[1:]).max() <= 1e-12 * 1.5
