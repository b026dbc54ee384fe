"""Host-side mirror of the reference's SparseMatrix and SparseVector (src/Math/SparseMatrix.h:34-81,
src/Math/SparseVector.h:8-43) over the caffe_matrix_* / caffe_vector_* entry points of
libcaffe_hip.so.

The method names, arguments and semantics are the reference's: ``setRow`` and ``setValue``
insert (an entry set again is overwritten; negative column and row indices are ignored, as
PETSc's MatSetValues ignores them), ``solve(b, x)`` starts from x = 0 and returns the
iteration count without raising when it does not converge (``last_report()`` tells), and
``vector(i)`` reads one entry.  The products and the solve run on the GPU in FP64."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

PETSC_DEFAULT = -2   # the constructor's default arguments (rtol 1e-5, abstol 1e-50, 10000 iterations)


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class SparseVector:
    def __init__(self, m: int | None = None, device: int = 0):
        self.lib = L.load()
        self.device = device
        self._h = C.c_void_p()
        self.m = 0
        if m is not None:
            self.allocate(m)

    def _check(self, rc, what):
        if rc != L.CAFFE_OK:
            msg = self.lib.caffe_vector_last_error(self._h).decode() if self._h else ""
            raise L.CaffeError(f"{what} failed ({rc}): {msg}")

    def allocate(self, m: int):
        self.deallocate()
        rc = self.lib.caffe_vector_create(self.device, int(m), C.byref(self._h))
        if rc != L.CAFFE_OK:
            raise L.CaffeError(f"caffe_vector_create failed ({rc})")
        self.m = int(m)

    def deallocate(self):
        if self._h:
            self.lib.caffe_vector_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.deallocate()
        except Exception:
            pass

    def setValue(self, i: int, value: float):
        self.set_values([i], [value])

    def set_values(self, idx, vals):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        self._check(self.lib.caffe_vector_set_values(self._h, idx.size, _ip(idx), _dp(vals)), "caffe_vector_set_values")

    def get_values(self, idx) -> np.ndarray:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros(idx.size)
        self._check(self.lib.caffe_vector_get_values(self._h, idx.size, _ip(idx), _dp(out)), "caffe_vector_get_values")
        return out

    def __call__(self, i: int) -> float:
        return float(self.get_values([i])[0])

    def zeroEntries(self):
        self._check(self.lib.caffe_vector_zero_entries(self._h), "caffe_vector_zero_entries")

    def _norm(self, t):
        out = C.c_double()
        self._check(self.lib.caffe_vector_norm(self._h, t, C.byref(out)), "caffe_vector_norm")
        return out.value

    def l1Norm(self):
        return self._norm(0)

    def l2Norm(self):
        return self._norm(1)

    def infNorm(self):
        return self._norm(2)

    def array(self) -> np.ndarray:
        return self.get_values(np.arange(self.m))


class SparseMatrix:
    def __init__(self, rToler: float = PETSC_DEFAULT, absToler: float = PETSC_DEFAULT, maxIters: int = PETSC_DEFAULT,
                 device: int = 0):
        self.lib = L.load()
        self.device = device
        self._tol = (rToler, absToler, maxIters)
        self._h = C.c_void_p()
        self.m = self.n = 0

    def _check(self, rc, what):
        if rc != L.CAFFE_OK:
            msg = self.lib.caffe_matrix_last_error(self._h).decode() if self._h else ""
            raise L.CaffeError(f"{what} failed ({rc}): {msg}")

    def allocate(self, m: int, n: int, nnz: int):
        self.deallocate()
        rt, at, mi = self._tol
        rc = self.lib.caffe_matrix_create(self.device, int(m), int(n), int(nnz), float(rt), float(at), int(mi),
                                          C.byref(self._h))
        if rc != L.CAFFE_OK:
            raise L.CaffeError(f"caffe_matrix_create failed ({rc})")
        self.m, self.n = int(m), int(n)

    def deallocate(self):
        if self._h:
            self.lib.caffe_matrix_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.deallocate()
        except Exception:
            pass

    def iLower(self):
        return 0

    def iUpper(self):
        return self.m

    def setValue(self, i: int, j: int, value: float):
        self._check(self.lib.caffe_matrix_set_value(self._h, int(i), int(j), float(value)), "caffe_matrix_set_value")

    def setRow(self, rowNo: int, nCols: int, colNos, values):
        self.set_rows([rowNo], np.array([0, nCols]), np.asarray(colNos)[:nCols], np.asarray(values)[:nCols])

    def set_rows(self, rows, rowptr, cols, vals):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        self._check(self.lib.caffe_matrix_set_rows(self._h, rows.size, _ip(rows), _ip(rowptr), _ip(cols), _dp(vals)),
                    "caffe_matrix_set_rows")

    def zeroEntries(self):
        self._check(self.lib.caffe_matrix_zero_entries(self._h), "caffe_matrix_zero_entries")

    def assemble(self):
        self._check(self.lib.caffe_matrix_assemble(self._h), "caffe_matrix_assemble")

    def solve(self, b: SparseVector, x: SparseVector) -> int:
        it = C.c_int()
        self._check(self.lib.caffe_matrix_solve(self._h, b._h, x._h, C.byref(it)), "caffe_matrix_solve")
        return it.value

    def last_report(self) -> dict:
        r = L.SolveReport()
        self._check(self.lib.caffe_matrix_last_report(self._h, C.byref(r)), "caffe_matrix_last_report")
        return r.as_dict()


def multiply(A: SparseMatrix, x: SparseVector, b: SparseVector):
    """b = A x (SparseMatrix.cpp:135-138)"""
    A._check(A.lib.caffe_matrix_multiply(A._h, x._h, b._h), "caffe_matrix_multiply")


def multiplyAdd(A: SparseMatrix, x1: SparseVector, x2: SparseVector, b: SparseVector):
    """b = A x1 + x2 (SparseMatrix.cpp:140-143)"""
    A._check(A.lib.caffe_matrix_multiply_add(A._h, x1._h, x2._h, b._h), "caffe_matrix_multiply_add")


def scale(alpha: float, obj):
    """scale(alpha, A) (SparseMatrix.cpp:145-148) or scale(alpha, v) (SparseVector.h:40)"""
    if isinstance(obj, SparseMatrix):
        obj._check(obj.lib.caffe_matrix_scale(float(alpha), obj._h), "caffe_matrix_scale")
    else:
        obj._check(obj.lib.caffe_vector_scale(float(alpha), obj._h), "caffe_vector_scale")
