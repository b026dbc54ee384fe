"""Why the learned stage (flexible CG) is not used on immersed-boundary p' systems.

The learned preconditioner acts inside flexible CG, which needs a symmetric positive
(semi)definite operator.  The reference's IbPiso replaces the p' row of every ghost cell by
its interpolation row (IbPiso.cpp:160-185: -1 on the diagonal, the trilinear weights beta
of the image point's eight cells elsewhere), while the fluid rows keep their symmetric
7-point coefficients: the operator becomes nonsymmetric.  These tests measure that on the
oracle's own assembled systems and show the consequence for a CG-type method:

- the fluid block alone is exactly symmetric, so all asymmetry is in the ghost rows;
- the whole IB p' matrix is far from symmetric;
- Jacobi-preconditioned CG diverges on it, while on the symmetric cavity p' system of the
  same kind (the systems the learned stage is trained and used on) it converges.
The IB p' systems are solved by the multigrid-preconditioned BiCGStab instead, whose
fine-level smoother solves the ghost rows exactly; the learned stage refuses them.
"""
import numpy as np
import pytest

scipy_sparse = pytest.importorskip("scipy.sparse")
from scipy.sparse import linalg as sla  # noqa: E402

from caffe_old_amd import case as cs  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402

from tests.helpers import threads  # noqa: E402


def _pcorr_system(o):
    rp, cols, vals = o.matrix(1)
    n = len(rp) - 1
    A = scipy_sparse.csr_matrix((vals, cols, rp), shape=(n, n))
    b = o.rhs(0)
    rid, st = o.imap()
    act = rid >= 0
    status = np.empty(n, dtype=int)
    status[rid[act]] = st[act]
    return A, b, status


def _jacobi_cg(A, b, maxiter):
    M = scipy_sparse.diags(1.0 / A.diagonal())
    x, _ = sla.cg(A, b, rtol=1e-10, atol=0.0, maxiter=maxiter, M=M)
    return np.linalg.norm(b - A @ x) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def ib_system():
    o = Oracle(cs.channel_ib_sphere(64, 16, 16, steps=1), rtol=1e-8, n_threads=threads())
    o.solve()  # one IbPiso step: the last p' system stays assembled
    return _pcorr_system(o)


def test_ib_pcorr_fluid_block_symmetric_ghost_rows_not(ib_system):
    A, _, st = ib_system
    ghost = np.where(st == 1)[0]
    fluid = np.where(st == 0)[0]
    assert len(ghost) >= 8
    Aff = A[fluid][:, fluid]
    assert scipy_sparse.linalg.norm(Aff - Aff.T) == 0.0
    asym = scipy_sparse.linalg.norm(A - A.T) / scipy_sparse.linalg.norm(A)
    print(f"IB p' system 64x16x16: {len(ghost)} ghost rows, ||A - A^T||_F / ||A||_F = {asym:.3f}")
    assert asym > 0.1
    # a ghost row: -1 (+ beta of its own cell) on the diagonal, the other betas off it, and
    # its cells' rows do not hold the mirrored entries
    g = ghost[0]
    row = A.getrow(g).toarray().ravel()
    col = A.getcol(g).toarray().ravel()
    off = [c for c in np.nonzero(row)[0] if c != g]
    assert any(abs(row[c] - col[c]) > 1e-6 for c in off)


def test_cg_diverges_on_ib_pcorr(ib_system):
    A, b, _ = ib_system
    rel = _jacobi_cg(A, b, 1000)
    print(f"Jacobi CG, 1000 iterations, IB p' system: relative residual {rel:.3e}")
    assert not rel < 1e-2


def test_cg_converges_on_cavity_pcorr():
    case = cs.lid_cavity(16, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.5 / 16,
                         n_pcorr=2, max_iterations=2)
    o = Oracle(case, n_threads=threads())
    o.solve()
    A, b, _ = _pcorr_system(o)
    assert scipy_sparse.linalg.norm(A - A.T) == 0.0
    b = b - b.mean()   # pure Neumann: the compatible right-hand side
    rel = _jacobi_cg(A, b, 1000)
    print(f"Jacobi CG, cavity 16^3 p' system: relative residual {rel:.3e}")
    assert rel < 1e-8
