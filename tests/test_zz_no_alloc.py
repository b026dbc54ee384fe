"""caffe_matrix_solve_noalloc (the solve whose BiCGStab work lives in the library's fixed
per-GPU scratch array) against caffe_matrix_solve on the same systems: two calls in a row, the
second on a smaller system, bit-identical solutions and iteration counts; the device's free
memory is the same before and after them (no allocation); too many rows and an unassembled
matrix are refused with an error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n, seed):
    """7-point diffusion rows with random positive coefficients on n^3 cells, columns -1 off the
    domain (dropped), and a random right-hand side"""
    rng = np.random.default_rng(seed)
    N = n ** 3
    q = np.arange(N)
    i, j, k = q % n, (q // n) % n, q // (n * n)
    cols, vals = [q], []
    diag = np.full(N, -0.05)
    for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        ii, jj, kk = i + di, j + dj, k + dk
        inside = (ii >= 0) & (jj >= 0) & (kk >= 0) & (ii < n) & (jj < n) & (kk < n)
        c = np.where(inside, (kk * n + jj) * n + ii, -1)
        a = rng.uniform(0.5, 1.5, N)
        cols.append(c)
        vals.append(a)
        diag -= a
    vals = [diag] + vals
    cols = np.stack(cols, axis=1).astype(np.int32).ravel()
    vals = np.stack(vals, axis=1).ravel()
    rowptr = (7 * np.arange(N + 1)).astype(np.int32)
    return N, rowptr, cols, vals, rng.normal(size=N)


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _make(n, seed):
    from caffe_old_amd.sparse import SparseMatrix, SparseVector
    N, rp, cols, vals, b = _system(n, seed)
    A = SparseMatrix(rToler=1e-10, maxIters=20000)
    A.allocate(N, N, 7)
    A.set_rows(np.arange(N), rp, cols, vals)
    bv = SparseVector(N)
    bv.set_values(np.arange(N), b)
    return A, bv, N


def _noalloc(A, b, x):
    it = C.c_int()
    rc = A.lib.caffe_matrix_solve_noalloc(A._h, b._h, x._h, C.byref(it))
    return rc, it.value


def test_noalloc_matches_allocating_solve():
    from caffe_old_amd import _lib as L
    from caffe_old_amd.sparse import SparseVector
    # load every kernel of the path once (a first launch may load code objects)
    Aw, bw, Nw = _make(6, 9)
    Aw.solve(bw, SparseVector(Nw))
    A1, b1, N1 = _make(40, 1)   # 64000 rows
    A2, b2, N2 = _make(24, 2)   # the second call: a smaller system
    A1.assemble()
    A2.assemble()
    x1n, x2n = SparseVector(N1), SparseVector(N2)
    before = _free_bytes()
    rc1, it1n = _noalloc(A1, b1, x1n)
    rc2, it2n = _noalloc(A2, b2, x2n)
    after = _free_bytes()
    assert rc1 == L.CAFFE_OK and rc2 == L.CAFFE_OK
    assert after == before, (before, after)
    r1n, r2n = x1n.array(), x2n.array()
    x1, x2 = SparseVector(N1), SparseVector(N2)
    it1 = A1.solve(b1, x1)
    it2 = A2.solve(b2, x2)
    rep = A2.last_report()
    print(f"noalloc: {it1n} and {it2n} iterations, allocating solve {it1} and {it2}, relres {rep['relres']:.2e}")
    assert (it1n, it2n) == (it1, it2)
    assert np.array_equal(r1n, x1.array()) and np.array_equal(r2n, x2.array())
    assert rep["converged"] == 1


def test_noalloc_refuses_what_it_cannot_do():
    from caffe_old_amd import _lib as L
    from caffe_old_amd.sparse import SparseMatrix, SparseVector
    A, b, N = _make(8, 3)   # rows set, not assembled
    rc, _ = _noalloc(A, b, SparseVector(N))
    assert rc == L.CAFFE_ERR_STATE and b"assemble" in A.lib.caffe_matrix_last_error(A._h)
    m = 262144 + 1   # CAFFE_NOALLOC_MAX_ROWS + 1
    big = SparseMatrix()
    big.allocate(m, m, 1)
    big.set_rows(np.arange(m), np.arange(m + 1), np.arange(m), np.ones(m))
    big.assemble()
    bb = SparseVector(m)
    rc, _ = _noalloc(big, bb, SparseVector(m))
    assert rc == L.CAFFE_ERR_ARG and b"CAFFE_NOALLOC_MAX_ROWS" in big.lib.caffe_matrix_last_error(big._h)
