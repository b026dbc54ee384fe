"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libref_geom.so")

KSP_PETSC_LIKE = 0
KSP_TIGHT = 1


class OcCase(C.Structure):
    _fields_ = [
        ("nI", C.c_int), ("nJ", C.c_int), ("nK", C.c_int),
        ("algo", C.c_int), ("steady", C.c_int),
        ("rho", C.c_double), ("mu", C.c_double), ("alphaU", C.c_double), ("alphaP", C.c_double),
        ("nInner", C.c_int), ("nPCorr", C.c_int),
        ("bcType", C.c_int * 6), ("bcVec", C.c_double * 18), ("bcVal", C.c_double * 6),
        ("uInit", C.c_double * 3), ("pInit", C.c_double),
        ("sphereR", C.c_double), ("sphereC", C.c_double * 3),
        ("kspKind", C.c_int), ("rtol", C.c_double), ("maxit", C.c_int),
        ("nBlocks", C.c_int), ("nThreads", C.c_int), ("nonOrth", C.c_int),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.oc_create.restype = C.c_void_p
        L.oc_create.argtypes = [C.POINTER(OcCase), dp]
        L.oc_destroy.argtypes = [C.c_void_p]
        for n in ("oc_solve", "oc_courant"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [C.c_void_p, C.c_double]
        for n in ("oc_compute_momentum", "oc_assemble_momentum"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_double]
        for n in ("oc_save_u0", "oc_compute_pcorr", "oc_assemble_pcorr", "oc_correct", "oc_rhie_chow",
                  "oc_compute_ib_field", "oc_compute_ib_coeffs", "oc_grad_u"):
            getattr(L, n).argtypes = [C.c_void_p]
        for n in ("oc_continuity", "oc_continuity_max"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [C.c_void_p]
        for n in ("oc_get", "oc_set", "oc_get_metric"):
            getattr(L, n).restype = C.c_long
            getattr(L, n).argtypes = [C.c_void_p, C.c_char_p, dp]
        L.oc_nrows.restype = C.c_int
        L.oc_nrows.argtypes = [C.c_void_p]
        L.oc_matrix_nnz.restype = C.c_long
        L.oc_matrix_nnz.argtypes = [C.c_void_p, C.c_int]
        L.oc_matrix_csr.argtypes = [C.c_void_p, C.c_int, ip, ip, dp]
        L.oc_rhs.argtypes = [C.c_void_p, C.c_int, dp]
        L.oc_imap.argtypes = [C.c_void_p, ip, ip]
        L.oc_stats.argtypes = [C.c_void_p, ip, ip, dp]
        L.oc_hex_volume.restype = C.c_double
        L.oc_hex_volume.argtypes = [dp]
        L.oc_hex_centroid.argtypes = [dp, dp]
        L.oc_quad_area.restype = C.c_double
        L.oc_quad_area.argtypes = [dp]
        L.oc_quad_centroid.argtypes = [dp, dp]
        L.oc_inside_hex.restype = C.c_int
        L.oc_inside_hex.argtypes = [dp, dp]
        L.oc_trilinear.argtypes = [dp, dp, dp]
        L.oc_transfinite_nodes.argtypes = [dp, C.c_int, C.c_int, C.c_int, dp]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own Geometry/Sphere/Array3D code (oracle/_ref), or None if not built."""
    if not os.path.exists(_REF):
        return None
    L = C.CDLL(_REF)
    dp = C.POINTER(C.c_double)
    L.ref_hex_volume.restype = C.c_double
    L.ref_hex_volume.argtypes = [dp]
    L.ref_hex_centroid.argtypes = [dp, dp]
    L.ref_quad_area.restype = C.c_double
    L.ref_quad_area.argtypes = [dp]
    L.ref_quad_centroid.argtypes = [dp, dp]
    L.ref_inside_hex.restype = C.c_int
    L.ref_inside_hex.argtypes = [dp, dp]
    L.ref_face_area_normal.argtypes = [dp, dp]
    L.ref_sphere_nearest.argtypes = [C.c_double, dp, dp, dp]
    L.ref_sphere_inside.restype = C.c_int
    L.ref_sphere_inside.argtypes = [C.c_double, dp, dp]
    L.ref_array3d_offset.restype = C.c_long
    L.ref_array3d_offset.argtypes = [C.c_int] * 6
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """One reference Solver instance (Simple / Piso / IbPiso) on a Case."""

    def __init__(self, case, ksp=KSP_TIGHT, rtol=1e-13, maxit=20000, n_blocks=1, n_threads=1):
        from caffe_old_amd.case import ALGOS, BC_TYPES
        self.case = case
        ni, nj, nk = case.shape
        c = OcCase()
        c.nI, c.nJ, c.nK = ni, nj, nk
        c.algo = ALGOS[case.algo]
        c.steady = 0 if case.time_accurate else 1
        c.rho, c.mu, c.alphaU, c.alphaP = case.rho, case.mu, case.alpha_u, case.alpha_p
        c.nInner, c.nPCorr = case.n_inner, case.n_pcorr
        for f, b in enumerate(case.boundaries):
            c.bcType[f] = BC_TYPES[b.type]
            for q in range(3):
                c.bcVec[3 * f + q] = float(b.refVector[q])
            c.bcVal[f] = float(b.refValue)
        for q in range(3):
            c.uInit[q] = case.u_init[q]
        c.pInit = case.p_init
        if case.sphere is not None:
            c.sphereR = case.sphere[0]
            for q in range(3):
                c.sphereC[q] = case.sphere[1][q]
        c.kspKind, c.rtol, c.maxit, c.nBlocks, c.nThreads = ksp, rtol, maxit, n_blocks, n_threads
        c.nonOrth = 1 if getattr(case, "non_orth", False) else 0
        self._c = c
        nodes = np.ascontiguousarray(case.nodes, dtype=np.float64)
        self.h = lib().oc_create(C.byref(c), _dp(nodes))
        self.shape = (ni, nj, nk)
        self.P = (ni + 2) * (nj + 2) * (nk + 2)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_destroy(self.h)
            self.h = None

    # -- stepping
    def solve(self, dt=None):
        return lib().oc_solve(self.h, self.case.time_step if dt is None else dt)

    def __getattr__(self, name):
        L = lib()
        fn = getattr(L, "oc_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda *a: fn(self.h, *a)

    # -- state in the padded device layout
    _NC = {"u": 3, "u0": 3, "h": 3, "gradP": 3, "gradPCorr": 3, "gradU": 9, "gradPhi": 3}

    def get(self, name):
        nc = self._NC.get(name, 1)
        out = np.zeros(nc * self.P)
        n = lib().oc_get(self.h, name.encode(), _dp(out))
        if n < 0:
            raise KeyError(name)
        return out.reshape(nc, self.P) if nc > 1 else out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if lib().oc_set(self.h, name.encode(), _dp(a)) < 0:
            raise KeyError(name)

    def metric(self, name):
        nc = 3 if name in ("xc", "SI", "SJ", "SK", "fcI", "fcJ", "fcK", "cE", "cW", "cN", "cS", "cT", "cB") else 1
        out = np.zeros(nc * self.P)
        if lib().oc_get_metric(self.h, name.encode(), _dp(out)) < 0:
            raise KeyError(name)
        return out.reshape(nc, self.P) if nc > 1 else out

    def matrix(self, which):
        """CSR of A_[which]; which 2 is the Coupled system (4 rows per cell)."""
        L = lib()
        n = L.oc_nrows(self.h) * (4 if which == 2 else 1)
        nnz = L.oc_matrix_nnz(self.h, which)
        rp = np.zeros(n + 1, dtype=np.int32)
        cols = np.zeros(max(nnz, 1), dtype=np.int32)
        vals = np.zeros(max(nnz, 1))
        L.oc_matrix_csr(self.h, which, _ip(rp), _ip(cols), _dp(vals))
        return rp, cols[:nnz], vals[:nnz]

    def rhs(self, vec):
        """b_[vec]; vec 3 is the Coupled right-hand side."""
        n = lib().oc_nrows(self.h) * (4 if vec == 3 else 1)
        out = np.zeros(n)
        lib().oc_rhs(self.h, vec, _dp(out))
        return out

    def imap(self):
        ni, nj, nk = self.shape
        rid = np.zeros(ni * nj * nk, dtype=np.int32)
        st = np.zeros(ni * nj * nk, dtype=np.int32)
        lib().oc_imap(self.h, _ip(rid), _ip(st))
        return rid, st

    def stats(self):
        a, b, r = C.c_int(), C.c_int(), C.c_double()
        lib().oc_stats(self.h, C.byref(a), C.byref(b), C.byref(r))
        return a.value, b.value, r.value


def interior(arr, shape):
    """View the interior cells of a padded array as (nK, nJ, nI) (or (nc, nK, nJ, nI))."""
    ni, nj, nk = shape
    a = np.asarray(arr)
    if a.ndim == 1:
        return a.reshape(nk + 2, nj + 2, ni + 2)[1:-1, 1:-1, 1:-1]
    return a.reshape(a.shape[0], nk + 2, nj + 2, ni + 2)[:, 1:-1, 1:-1, 1:-1]
