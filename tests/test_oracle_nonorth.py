"""Closed-form pin of the oracle's corrected (non-orthogonal) mode: the momentum term
mu_f (grad u)_f . c_f of Simple.cpp:153-158 (oracle/caffe_oracle.cpp: assembleMomentum with
gradUFaces).

The mesh: rows of cells with different shear, x = X_i + sigma_j, y = Y_j (stretched),
z = Z_k, so every cell is a parallelogram extruded in z and every face is planar.  Face
centroids, area normals, cell centroids and volumes are then closed forms of X, Y, Z,
sigma, and so is c = S - r (S.S)/(S.r) (r cell to cell inside, cell to face on the
boundary, HexaFvmMesh.cpp:218-294, 879-884); because the shear changes from row to row,
the c of a cell's north and south faces do not cancel.

With a frozen state (no face flux, u = 0 on cells and patches, grad p = 0, steady,
alpha = 1) and a given grad u per cell, the momentum right-hand side is that term alone:
  b = sum_f mu (grad u)_f c_f,
with (grad u)_f = g G_P + (1 - g) G_nb on an interior face (g = V_nb / (V_P + V_nb)) and
G_P on a domain face.  In reference mode the faces of grad u are never assigned and b = 0.
A wrong sign of the term, a c with the wrong sign on the low faces, a transposed tensor,
swapped weights or the wrong boundary value fails the comparison."""
import numpy as np
import pytest

from caffe_old_amd import case as cs
from oracle.pyoracle import Oracle, interior
from tests.helpers import oracle_rhs_padded

pytestmark = pytest.mark.usefixtures("oracle_built")


def _mesh(ni=5, nj=6, nk=3):
    X = np.concatenate([[0.0], np.cumsum(1.2 ** np.arange(ni))])
    X /= X[-1]
    Y = np.concatenate([[0.0], np.cumsum(0.8 ** np.arange(nj))])
    Y /= Y[-1]
    Z = np.linspace(0.0, 0.6, nk + 1)
    sig = 0.13 * np.sin(2.1 * np.arange(nj + 1)) + 0.05 * np.arange(nj + 1) ** 1.5 / nj
    nodes = np.zeros((nk + 1, nj + 1, ni + 1, 3))
    nodes[..., 0] = X[None, None, :] + sig[None, :, None]
    nodes[..., 1] = Y[None, :, None]
    nodes[..., 2] = Z[:, None, None]
    return nodes, X, Y, Z, sig


def _geometry(X, Y, Z, sig):
    """per cell (k, j, i): volume, centroid; per face direction the six (S outward, r) pairs"""
    ni, nj, nk = len(X) - 1, len(Y) - 1, len(Z) - 1
    dX, h, dZ = np.diff(X), np.diff(Y), np.diff(Z)
    Xc, Yc, Zc = 0.5 * (X[1:] + X[:-1]), 0.5 * (Y[1:] + Y[:-1]), 0.5 * (Z[1:] + Z[:-1])
    sc = 0.5 * (sig[1:] + sig[:-1])
    K, J, I = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    vol = dX[I] * h[J] * dZ[K]
    xc = np.stack([Xc[I] + sc[J], Yc[J], Zc[K]], axis=-1)
    z3 = np.zeros_like(vol)
    # outward area normals
    SE = np.stack([h[J] * dZ[K], -(sig[J + 1] - sig[J]) * dZ[K], z3], axis=-1)
    SN = np.stack([z3, dX[I] * dZ[K], z3], axis=-1)
    ST = np.stack([z3, z3, dX[I] * h[J]], axis=-1)
    # face centroids
    fE = np.stack([X[I + 1] + sc[J], Yc[J], Zc[K]], axis=-1)
    fW = np.stack([X[I] + sc[J], Yc[J], Zc[K]], axis=-1)
    fN = np.stack([Xc[I] + sig[J + 1], Y[J + 1], Zc[K]], axis=-1)
    fS = np.stack([Xc[I] + sig[J], Y[J], Zc[K]], axis=-1)
    fT = np.stack([Xc[I] + sc[J], Yc[J], Z[K + 1]], axis=-1)
    fB = np.stack([Xc[I] + sc[J], Yc[J], Z[K]], axis=-1)

    def nb(axis, step):
        """centroid of the neighbour one step along axis (NaN beyond the domain edge)"""
        sl = [slice(None)] * 3
        out = np.full_like(xc, np.nan)
        src = [slice(None)] * 3
        ax = 2 - axis
        if step > 0:
            sl[ax], src[ax] = slice(0, -1), slice(1, None)
        else:
            sl[ax], src[ax] = slice(1, None), slice(0, -1)
        out[tuple(sl)] = xc[tuple(src)]
        return out

    # the outward normals of the low faces: the I-faces of a row are parallel copies of each
    # other, the J-faces lie in planes y = const and have area dX dZ, the K-faces in z = const
    faces = []
    for S, fc, axis, step in ((SE, fE, 0, 1), (-SE, fW, 0, -1), (SN, fN, 1, 1), (-SN, fS, 1, -1),
                              (ST, fT, 2, 1), (-ST, fB, 2, -1)):
        other = nb(axis, step)
        r = np.where(np.isnan(other), fc - xc, other - xc)
        faces.append((S, r, axis, step))
    return vol, xc, faces


def _c(S, r):
    ss = np.einsum("...i,...i", S, S)
    sr = np.einsum("...i,...i", S, r)
    return S - r * (ss / sr)[..., None]


@pytest.mark.parametrize("corrected", [True, False])
def test_nonorth_term_closed_form(corrected):
    nodes, X, Y, Z, sig = _mesh()
    b = [cs.Boundary("wall") for _ in range(6)]
    mu = 0.037
    case = cs.Case(nodes=nodes, algo="simple", rho=1.0, mu=mu, alpha_u=1.0, alpha_p=1.0, boundaries=b, max_iterations=2,
                   non_orth=corrected)
    o = Oracle(case)
    P = o.P
    ni, nj, nk = case.shape
    for n in ("mI", "mJ", "mK"):
        o.set(n, np.zeros(P))
    o.set("u", np.zeros((3, P)))
    o.set("gradP", np.zeros((3, P)))
    rng = np.random.default_rng(4)
    Gc = rng.normal(size=(nk, nj, ni, 9))
    gu = np.zeros((9, P))
    for q in range(9):
        interior(gu[q], case.shape)[...] = Gc[..., q]
    o.set("gradU", gu)
    o.assemble_momentum(1.0)
    got = np.stack([interior(oracle_rhs_padded(o, c, P), case.shape) for c in range(3)], axis=-1)
    if not corrected:
        assert np.abs(got).max() == 0.0
        return
    vol, xc, faces = _geometry(X, Y, Z, sig)
    expect = np.zeros((nk, nj, ni, 3))
    for S, r, axis, step in faces:
        c = _c(S, r)
        ax = 2 - axis
        # face tensor: interpolated with the volume weight inside, the cell's own at the edge
        Gn = np.roll(Gc, -step, axis=ax)
        Vn = np.roll(vol, -step, axis=ax)
        g = (Vn / (Vn + vol))[..., None]
        Tf = g * Gc + (1 - g) * Gn
        edge = [slice(None)] * 3
        edge[ax] = -1 if step > 0 else 0
        Tf[tuple(edge)] = Gc[tuple(edge)]
        T = Tf.reshape(nk, nj, ni, 3, 3)
        expect += mu * np.einsum("...ij,...j->...i", T, c)
    # the north/south c of a cell do not cancel on this mesh, so the term is not trivially zero
    assert np.abs(expect).max() > 1e-3 * mu
    assert np.abs(got - expect).max() <= 1e-12 * np.abs(expect).max()
