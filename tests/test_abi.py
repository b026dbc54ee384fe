"""CPU checks of the drop-in boundary: libcaffe_hip.so loads without a GPU and exports
every function include/caffe_hip.h declares; case validation on the host side."""
import ctypes
import os
import re

import pytest

from caffe_old_amd import _lib
from caffe_old_amd import case as cs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "caffe_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|const char\*)\s+(caffe_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == declared


def test_product_fails_loudly_without_library(tmp_path):
    with pytest.raises(_lib.CaffeError):
        _lib._lib = None
        try:
            _lib.load(str(tmp_path / "missing.so"))
        finally:
            _lib._lib = None


def test_case_validation():
    with pytest.raises(ValueError):
        cs.Boundary("slip")  # FlowBoundaryConditions.cpp:41-42: unknown type is an error
    c = cs.cavity_simple_32()
    assert c.shape == (32, 32, 32) and c.n_solves == 200  # RunControl.cpp:53-63 off-by-one
    p = cs.cavity_piso_128(16, steps=3)
    assert p.time_step == 0.5 / 16 and p.n_pcorr == 2 and p.n_solves == 3
    ch = cs.channel_ib_sphere(32, 8, 8, steps=2)
    assert ch.shape == (32, 8, 8) and ch.time_step == 0.5 * 4.0 / 32
    with pytest.raises(ValueError):
        cs.Case(nodes=ch.nodes, algo="ibpiso")


def test_perturbation_keeps_boundary_nodes():
    n = cs.cartesian_nodes(1, 1, 1, 6, 6, 6)
    q = cs.perturb_interior(n, 0.15, 1 / 6, seed=0)
    assert (q[0] == n[0]).all() and (q[-1] == n[-1]).all()
    assert (q[:, 0] == n[:, 0]).all() and (q[:, :, -1] == n[:, :, -1]).all()
    d = (q - n)[1:-1, 1:-1, 1:-1]
    assert 0 < abs(d).max() <= 0.15 / 6
    # counter-based: the same node draws the same number in any sub-range
    assert (cs.perturb_interior(n, 0.15, 1 / 6, seed=0) == q).all()
    assert not (cs.perturb_interior(n, 0.15, 1 / 6, seed=1) == q).all()


def test_philox_known_answer():
    """Philox-4x32-10 known-answer vectors (Random123 kat_vectors: counter 0, key 0 and
    counter/key all ones)."""
    import numpy as np
    z = cs.philox4x32(np.zeros((1, 4), dtype=np.uint32), (0, 0))[0]
    assert [hex(int(v)) for v in z] == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    f = cs.philox4x32(np.full((1, 4), 0xFFFFFFFF, dtype=np.uint32), (0xFFFFFFFF, 0xFFFFFFFF))[0]
    assert [hex(int(v)) for v in f] == ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]


def test_solve_mode_constants_match_header():
    txt = open(os.path.join(ROOT, "include", "caffe_hip.h")).read()
    modes = dict((k, int(v)) for k, v in re.findall(r"CAFFE_SOLVE_(\w+)\s*=\s*(\d+)", txt))
    assert modes == {"PLAIN_FP64": _lib.SOLVE_PLAIN_FP64, "IR_FP32": _lib.SOLVE_IR_FP32,
                     "LEARNED_FCG": _lib.SOLVE_LEARNED_FCG, "MG": _lib.SOLVE_MG, "SGS_FP64": _lib.SOLVE_SGS_FP64}


def test_cpp_wrapper_header_compiles(tmp_path):
    """include/caffe_hip.hpp (the reference-shaped SparseMatrix / SparseVector / Solver classes
    over the C ABI) compiles as C++11 with g++ alone, and the drop-in test program built with
    it links against the in-tree library."""
    import subprocess
    src = tmp_path / "t.cpp"
    src.write_text('#include "caffe_hip.hpp"\nint main() { caffe_gpu::SparseMatrix A; (void)A; return 0; }\n')
    subprocess.check_call(["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           str(src)])
    exe = os.path.join(ROOT, "caffe_old_amd", "csrc", "_build", "drop_in")
    assert os.path.exists(exe), "run __graft_entry__.build()"
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libcaffe_hip.so" in out and "not found" not in out, out
