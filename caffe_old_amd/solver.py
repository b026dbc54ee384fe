"""Host-side mirror of CAFFE's Solver interface for the SIMPLE / PISO / IbPiso path.

``Simple``, ``Piso`` and ``IbPiso`` keep the reference's shape: construct from a case
(the reference constructs from ``Input`` and ``HexaFvmMesh``, Simple.cpp:34-74), call
``solve(dt)`` once per RunControl iteration (Solver.h:20) and read the continuity value
it returns, then ``displayUpdateMessage()`` (Simple.cpp:93-100, Piso.cpp:31-38).  The
work runs on one MI355X through the C ABI of libcaffe_hip.so; fields stay resident in
HBM and are read back only on request.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .case import ALGOS, BC_TYPES, Case


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """One GPU-resident solver instance (the reference's Solver subclasses)."""

    NCOMP = {"u": 3, "u0": 3, "h": 3, "gradP": 3, "gradPCorr": 3, "gradU": 9, "gradPhi": 3}

    def __init__(self, case: Case, device: int = 0, uniform: bool | None = None):
        self.lib = L.load()
        self.case = case
        self._h = C.c_void_p()
        self._check(self.lib.caffe_ctx_create(device, C.byref(self._h)), "caffe_ctx_create")
        ni, nj, nk = case.shape
        cs = L.CaffeCase()
        cs.nI, cs.nJ, cs.nK = ni, nj, nk
        cs.algo = ALGOS[case.algo]
        cs.steady = 0 if case.time_accurate else 1
        cs.rho, cs.mu, cs.alphaU, cs.alphaP = case.rho, case.mu, case.alpha_u, case.alpha_p
        cs.nInner, cs.nPCorr = case.n_inner, case.n_pcorr
        for f, b in enumerate(case.boundaries):
            cs.bcType[f] = BC_TYPES[b.type]
            for q in range(3):
                cs.bcVec[3 * f + q] = float(b.refVector[q])
            cs.bcVal[f] = float(b.refValue)
        for q in range(3):
            cs.uInit[q] = case.u_init[q]
        cs.pInit = case.p_init
        if case.sphere is not None:
            cs.sphereR = case.sphere[0]
            for q in range(3):
                cs.sphereC[q] = case.sphere[1][q]
        if uniform is None:
            uniform = _is_uniform_box(case.nodes)
        cs.uniform = 1 if uniform else 0
        cs.nonOrth = 1 if getattr(case, "non_orth", False) else 0
        self.uniform = bool(uniform)
        nodes = np.ascontiguousarray(case.nodes, dtype=np.float64)
        self._check(self.lib.caffe_case_setup(self._h, C.byref(cs), _dp(nodes)), "caffe_case_setup")
        self.shape = (ni, nj, nk)
        self.P = int(self.lib.caffe_padded_size(self._h))
        self.continuityError = 0.0
        self.courantNumber = 0.0

    # -------------------------------------------------------------- plumbing
    def _check(self, rc, what):
        if rc != L.CAFFE_OK:
            msg = self.lib.caffe_last_error(self._h).decode() if self._h else ""
            raise L.CaffeError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            if getattr(self, "_net", None) is not None:
                self.lib.caffe_ctx_attach_net(self._h, None)
                self._net = None
            self.lib.caffe_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def attach_net(self, net):
        """Use a trained ``caffe_old_amd.learned.PressureNet`` in SOLVE_LEARNED_FCG solves
        (None detaches).  The solver keeps a reference so the network outlives it."""
        self._check(self.lib.caffe_ctx_attach_net(self._h, net._h if net is not None else None), "caffe_ctx_attach_net")
        self._net = net

    def set_fcg_coarse(self, mode: int):
        """coarse solve of the learned two-grid cycle: 0 the network, 1 a multigrid V-cycle
        (the ablation)"""
        self._check(self.lib.caffe_ctx_set_fcg_coarse(self._h, int(mode)), "caffe_ctx_set_fcg_coarse")

    def set_fcg_post_sweeps(self, sweeps: int):
        """damped Jacobi sweeps after the learned cycle's coarse correction (1-4, default 3)"""
        self._check(self.lib.caffe_ctx_set_fcg_post_sweeps(self._h, int(sweeps)), "caffe_ctx_set_fcg_post_sweeps")

    def set_fcg_guess(self, on: bool):
        """learned and multigrid p' solves start from the best combination of the same
        corrector's solutions in the two previous steps (True, default) or from x = 0"""
        self._check(self.lib.caffe_ctx_set_fcg_guess(self._h, 1 if on else 0), "caffe_ctx_set_fcg_guess")

    def set_solve_opts(self, which: str, mode: int = L.SOLVE_PLAIN_FP64, rtol: float = 1e-8, maxit: int = 50000,
                       inner_rtol: float = 1e-4):
        o = L.SolveOpts(mode, rtol, maxit, inner_rtol)
        self._check(self.lib.caffe_set_solve_opts(self._h, 0 if which == "momentum" else 1, C.byref(o)),
                    "caffe_set_solve_opts")

    # -------------------------------------------------------------- the reference interface
    def solve(self, dt: float | None = None) -> float:
        """Solver::solve(timeStep): one outer iteration / time step; returns the continuity
        value of Simple::computeContinuityError (Simple.cpp:358-380) and keeps the Courant
        number the step computed (Simple.cpp:88, Piso.cpp:26)."""
        e, co = C.c_double(), C.c_double()
        self._check(self.lib.caffe_solve(self._h, self.case.time_step if dt is None else dt, C.byref(e)), "caffe_solve")
        self._check(self.lib.caffe_last_courant(self._h, C.byref(co)), "caffe_last_courant")
        self.continuityError = e.value
        self.courantNumber = co.value
        return e.value

    def displayUpdateMessage(self) -> str:
        m, p = self.reports()
        name = type(self).__name__
        msg = (f"{name}: completed iteration.\n"
               f"{name}: BiCGStab iterations  : {p['iters']}\n"
               f"{name}: Max continuity error : {self.continuityError:.6f}\n"
               f"{name}: Max Courant number   : {self.courantNumber:.6f}")
        print(msg)
        return msg

    # -------------------------------------------------------------- sub-steps
    def save_u0(self):
        self._check(self.lib.caffe_save_u0(self._h), "caffe_save_u0")

    def momentum_assemble(self, dt):
        self._check(self.lib.caffe_momentum_assemble(self._h, dt), "caffe_momentum_assemble")

    def momentum_solve(self):
        r = L.SolveReport()
        self._check(self.lib.caffe_momentum_solve(self._h, C.byref(r)), "caffe_momentum_solve")
        return r.as_dict()

    def rhie_chow(self):
        self._check(self.lib.caffe_rhiechow(self._h), "caffe_rhiechow")

    def pcorr_assemble(self):
        self._check(self.lib.caffe_pcorr_assemble(self._h), "caffe_pcorr_assemble")

    def pcorr_solve(self, mode=L.SOLVE_PLAIN_FP64, rtol=1e-8, maxit=50000, inner_rtol=1e-4):
        o = L.SolveOpts(mode, rtol, maxit, inner_rtol)
        r = L.SolveReport()
        self._check(self.lib.caffe_pcorr_solve(self._h, C.byref(o), C.byref(r)), "caffe_pcorr_solve")
        return r.as_dict()

    def correct(self):
        self._check(self.lib.caffe_correct(self._h), "caffe_correct")

    def ib_field(self):
        self._check(self.lib.caffe_ib_field(self._h), "caffe_ib_field")

    def ib_coeffs(self):
        self._check(self.lib.caffe_ib_coeffs(self._h), "caffe_ib_coeffs")

    def grad_u(self):
        self._check(self.lib.caffe_grad_u(self._h), "caffe_grad_u")

    def step_report(self, dt=None):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self._check(self.lib.caffe_step_report(self._h, self.case.time_step if dt is None else dt, C.byref(a),
                                               C.byref(b), C.byref(c)), "caffe_step_report")
        self.courantNumber = c.value
        return {"continuity": a.value, "continuity_max": b.value, "courant": c.value}

    def reports(self):
        m, p = L.SolveReport(), L.SolveReport()
        self._check(self.lib.caffe_last_reports(self._h, C.byref(m), C.byref(p)), "caffe_last_reports")
        return m.as_dict(), p.as_dict()

    def time_kernel(self, which: str, reps: int = 50):
        ms, by = C.c_double(), C.c_double()
        self._check(self.lib.caffe_time_kernel(self._h, which.encode(), reps, C.byref(ms), C.byref(by)), "caffe_time_kernel")
        return ms.value, by.value

    def kernel_probe(self, which: str, enable: bool = True):
        """Start (or stop) live event timing of one kernel inside the following solves."""
        self._check(self.lib.caffe_kernel_probe(self._h, which.encode(), 1 if enable else 0), "caffe_kernel_probe")

    def kernel_probe_read(self):
        """(average ms per launch that ran, number of such launches, algorithmic bytes per launch)"""
        ms, n, by = C.c_double(), C.c_long(), C.c_double()
        self._check(self.lib.caffe_kernel_probe_read(self._h, C.byref(ms), C.byref(n), C.byref(by)),
                    "caffe_kernel_probe_read")
        return ms.value, n.value, by.value

    def device_bytes(self) -> int:
        """Device memory held by the context (an attached network holds its own)."""
        return int(self.lib.caffe_device_bytes(self._h))

    def sync(self):
        self._check(self.lib.caffe_sync(self._h), "caffe_sync")

    # -------------------------------------------------------------- state (padded layout)
    def get(self, name: str) -> np.ndarray:
        nc = self.NCOMP.get(name, 1)
        out = np.zeros(nc * self.P)
        self._check(self.lib.caffe_field_download(self._h, name.encode(), _dp(out), out.size), f"download {name}")
        return out.reshape(nc, self.P) if nc > 1 else out

    def set(self, name: str, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._check(self.lib.caffe_field_upload(self._h, name.encode(), _dp(a), a.size), f"upload {name}")

    def metric(self, name: str) -> np.ndarray:
        nc = 3 if name in ("xc", "SI", "SJ", "SK", "fcI", "fcJ", "fcK") else 1
        out = np.zeros(nc * self.P)
        self._check(self.lib.caffe_metric_download(self._h, name.encode(), _dp(out), out.size), f"metric {name}")
        return out.reshape(nc, self.P) if nc > 1 else out

    def matrix(self, which: int) -> np.ndarray:
        """which 0: momentum band (7, P); 1: p' storage (4, P); 2: Coupled planes (56, P)."""
        n = {0: 7, 1: 4, 2: 56}[which] * self.P
        out = np.zeros(n)
        self._check(self.lib.caffe_matrix_download(self._h, which, _dp(out), n), "caffe_matrix_download")
        return out.reshape(-1, self.P)

    def rhs(self, which: int) -> np.ndarray:
        """which 0: momentum (3, P); 1: p' (P,); 2: Coupled (4, P)."""
        n = {0: 3, 1: 1, 2: 4}[which] * self.P
        out = np.zeros(n)
        self._check(self.lib.caffe_rhs_download(self._h, which, _dp(out), n), "caffe_rhs_download")
        return out.reshape(-1, self.P) if which != 1 else out

    def ghost_rows(self):
        n = C.c_int()
        self._check(self.lib.caffe_ghost_rows(self._h, C.byref(n), None, None, None, None), "caffe_ghost_rows")
        m = n.value
        cells = np.zeros(max(m, 1), dtype=np.int64)
        cols = np.zeros(9 * max(m, 1), dtype=np.int64)
        vm = np.zeros(9 * max(m, 1))
        vp = np.zeros(9 * max(m, 1))
        lp = C.POINTER(C.c_long)
        self._check(self.lib.caffe_ghost_rows(self._h, C.byref(n), cells.ctypes.data_as(lp), cols.ctypes.data_as(lp),
                                              _dp(vm), _dp(vp)), "caffe_ghost_rows")
        return cells[:m], cols[:9 * m].reshape(m, 9), vm[:9 * m].reshape(m, 9), vp[:9 * m].reshape(m, 9)


class Simple(Solver):
    """SIMPLE (Simple.cpp): one momentum solve and one pressure correction per iteration."""


class Piso(Solver):
    """PISO as the reference implements it (Piso.cpp:12-29): one momentum solve, then
    numberOfPressureCorrections x (pressure correction + correct)."""


class IbPiso(Solver):
    """PISO with a ghost-cell immersed sphere (IbPiso.cpp)."""


class Diffusion(Solver):
    """Diffusion (Diffusion.cpp:33-131): one steady or unsteady diffusion solve of phi per
    iteration; ``solve`` returns ||A x - b||_inf.  Fields "phi" and "gradPhi"; the solve
    options are set as those of the pressure correction (``set_solve_opts("pcorr", ...)``)."""

    def displayUpdateMessage(self) -> str:
        m, p = self.reports()
        msg = (f"Diffusion: completed iteration.\n"
               f"Diffusion: solver iterations : {p['iters']}\n"
               f"Diffusion: residual (inf)    : {self.continuityError:.6e}")
        print(msg)
        return msg


class Coupled(Solver):
    """Coupled (Coupled.cpp:38-338): u, v, w and p in one block system per inner iteration,
    solved in FP64 (Jacobi BiCGStab over the joint vector); ``solve`` returns the max over
    cells of |sum of face mass flows| / V.  The solve options are set as those of the
    pressure correction (``set_solve_opts("pcorr", rtol=..., maxit=...)``); the mode is not
    used.  Initial u and p are zero whatever the case holds (the reference reads no initial
    conditions for this solver)."""

    def displayUpdateMessage(self) -> str:
        m, p = self.reports()
        msg = ("Solver: completed iteration.\n"
               f"Solver: BiCGStab iterations: {p['iters']}")
        print(msg)
        return msg


def make_solver(case: Case, device: int = 0, uniform: bool | None = None) -> Solver:
    cls = {"simple": Simple, "piso": Piso, "ibpiso": IbPiso, "diffusion": Diffusion, "coupled": Coupled}[case.algo]
    return cls(case, device=device, uniform=uniform)


def _is_uniform_box(nodes: np.ndarray) -> bool:
    """True when the nodes are StructuredMesh::initializeCartesianMesh's box at the origin."""
    nk1, nj1, ni1, _ = nodes.shape
    lx, ly, lz = nodes[0, 0, -1, 0], nodes[0, -1, 0, 1], nodes[-1, 0, 0, 2]
    hx, hy, hz = lx / (ni1 - 1), ly / (nj1 - 1), lz / (nk1 - 1)
    k, j, i = np.meshgrid(np.arange(nk1), np.arange(nj1), np.arange(ni1), indexing="ij")
    ok = (np.abs(nodes[..., 0] - i * hx) <= 1e-12 * hx).all()
    ok &= (np.abs(nodes[..., 1] - j * hy) <= 1e-12 * hy).all()
    ok &= (np.abs(nodes[..., 2] - k * hz) <= 1e-12 * hz).all()
    return bool(ok)
