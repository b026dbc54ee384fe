"""The learned multilevel preconditioner of the pressure-correction solve.

``PressureNet`` owns a device-resident network (FP32 master parameters, BF16 packed
copies for the MFMA convolutions), its Adam state, and a device-resident training set of
p' systems collected from solver contexts.  Training and inference run entirely in the
HIP library (csrc/net.hip, csrc/learned.hip); this module only moves parameters and
small test arrays across the C ABI.

The cycle it defines for a residual r on the fine grid (dimensions divisible by 16):
    z1 = w D^-1 r;  c = net(P^T D^-1 (r - A z1));  y = z1 + P c;  z = y + w D^-1 (r - A y)
(w = 0.8, P = cell-centred trilinear prolongation).  Solves with
``SOLVE_LEARNED_FCG`` use it inside flexible CG; no counterpart exists in the reference,
whose pressure solve is PETSc BiCGStab + SOR (SparseMatrix.cpp:86-106).
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _lib as L

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def param_count() -> int:
    return int(L.load().caffe_net_param_count())


class PressureNet:
    def __init__(self, device: int = 0, seed: int = 0):
        self.lib = L.load()
        self._h = C.c_void_p()
        rc = self.lib.caffe_net_create(device, seed, C.byref(self._h))
        if rc != L.CAFFE_OK:
            raise L.CaffeError(f"caffe_net_create failed ({rc})")
        self.n_params = int(self.lib.caffe_net_param_count())
        self.post_sweeps = 1

    def _check(self, rc, what):
        if rc != L.CAFFE_OK:
            msg = self.lib.caffe_net_last_error(self._h).decode() if self._h else ""
            raise L.CaffeError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.caffe_net_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- parameters
    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        self._check(self.lib.caffe_net_get_params(self._h, _fp(out), out.size), "caffe_net_get_params")
        return out

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        self._check(self.lib.caffe_net_set_params(self._h, _fp(t), t.size), "caffe_net_set_params")

    def save(self, path: str, meta: dict | None = None):
        np.save(path, self.get_params())
        if meta is not None:
            with open(os.path.splitext(path)[0] + ".json", "w") as f:
                json.dump(meta, f, indent=1)

    def load(self, path: str):
        self.set_params(np.load(path, allow_pickle=False))
        return self

    # ---------------------------------------------------------- kernels on host arrays (tests)
    def forward(self, v: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float32)
        K, J, I = v.shape
        out = np.empty_like(v)
        self._check(self.lib.caffe_net_forward(self._h, I, J, K, _fp(v), _fp(out)), "caffe_net_forward")
        return out

    def backward(self, v: np.ndarray, gout: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float32)
        g = np.ascontiguousarray(gout, dtype=np.float32)
        K, J, I = v.shape
        out = np.empty(self.n_params, dtype=np.float32)
        self._check(self.lib.caffe_net_backward(self._h, I, J, K, _fp(v), _fp(g), _fp(out)), "caffe_net_backward")
        return out

    def set_post_sweeps(self, sweeps: int):
        """post-smoothing sweeps of the cycle in precond, loss_grad, train and eval (1-4)"""
        self._check(self.lib.caffe_net_set_post_sweeps(self._h, int(sweeps)), "caffe_net_set_post_sweeps")
        self.post_sweeps = int(sweeps)

    def precond(self, shape, A: np.ndarray, r: np.ndarray) -> np.ndarray:
        nI, nJ, nK = shape
        A = np.ascontiguousarray(A, dtype=np.float32)
        r = np.ascontiguousarray(r, dtype=np.float32)
        z = np.empty_like(r)
        self._check(self.lib.caffe_net_precond(self._h, nI, nJ, nK, _fp(A), _fp(r), _fp(z)), "caffe_net_precond")
        return z

    def loss_grad(self, shape, A, b, x, grad: bool = True):
        nI, nJ, nK = shape
        A = np.ascontiguousarray(A, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        loss = C.c_double()
        g = np.empty(self.n_params, dtype=np.float32) if grad else None
        self._check(self.lib.caffe_net_loss_grad(self._h, nI, nJ, nK, _fp(A), _fp(b), _fp(x), C.byref(loss),
                                                 _fp(g) if grad else None), "caffe_net_loss_grad")
        return (loss.value, g) if grad else loss.value

    # ---------------------------------------------------------- training set and training
    def collect(self, solver) -> int:
        """Append the solver's current p' system (after its pressure-correction solve)."""
        n = C.c_long()
        self._check(self.lib.caffe_net_collect(self._h, solver._h, C.byref(n)), "caffe_net_collect")
        return n.value

    def clear(self):
        self._check(self.lib.caffe_net_clear(self._h), "caffe_net_clear")

    @property
    def dataset_size(self) -> int:
        return int(self.lib.caffe_net_dataset_size(self._h))

    def train(self, steps: int, batch: int = 4, lr: float = 1e-3, seed: int = 0, synthetic: float = 0.5) -> np.ndarray:
        """`steps` Adam steps at rate lr; each batch entry is a stored pair or, with
        probability `synthetic`, an exact synthetic pair (cosine modes, b = A x)."""
        losses = np.zeros(steps, dtype=np.float32)
        self._check(self.lib.caffe_net_train(self._h, steps, batch, lr, synthetic, seed, _fp(losses)), "caffe_net_train")
        return losses

    def eval(self) -> np.ndarray:
        n = self.dataset_size
        out = np.zeros(n, dtype=np.float32)
        self._check(self.lib.caffe_net_eval(self._h, _fp(out), n), "caffe_net_eval")
        return out


def trained_weights_path(n: int = 128) -> str:
    return os.path.join(WEIGHTS_DIR, f"pcorr_net_cavity{n}.npy")
