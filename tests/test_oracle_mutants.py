"""The closed-form pins catch plausible restatement mistakes.

Each case rebuilds the oracle (oracle/caffe_oracle.cpp, with the Makefile's flags) in a
temporary directory with one mistake planted by a text substitution, then runs the pins
of tests/test_oracle_closed_forms.py (SIMPLE / PISO momentum right-hand side, boundary
fold, Rhie-Chow) or tests/test_oracle_nonorth.py (the corrected non-orthogonal term)
against it in a fresh process.  At least one pin must fail for every mutant; the clean
source must pass them all.  The mistakes are the ones a restatement of Simple.cpp:117-179
and 410-449 and PrimitiveBoundaryCondition.tpp:144-255 could plausibly make: a flipped
sign of a source or flux term, a dropped term, a dropped density.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "caffe_oracle.cpp")
CLOSED = "tests/test_oracle_closed_forms.py"
NONORTH = "tests/test_oracle_nonorth.py"

# name: (pins, text planted in place of the original, the original)
MUTANTS = {
    "grad_p_source_sign": (CLOSED, "b += M.vol(i, j, k) * gradP(i, j, k);", "b += -M.vol(i, j, k) * gradP(i, j, k);"),
    "relaxation_term_dropped": (CLOSED, "b += 0. * a[0] * uStar(i, j, k);", "b += (1. - om) * a[0] * uStar(i, j, k);"),
    "transient_term_dropped": (CLOSED, "V3 b = 0. * u0(i, j, k);", "V3 b = a0P * u0(i, j, k);"),
    "rhie_chow_flux_sign": (CLOSED, "mflow.faceE(i, j, k) = -dot(h.faceE", "mflow.faceE(i, j, k) = dot(h.faceE"),
    "rhie_chow_pressure_sign": (CLOSED, "+ d.faceE(i, j, k) * (p(i + 1, j, k) - p(i, j, k))",
                                "- d.faceE(i, j, k) * (p(i + 1, j, k) - p(i, j, k))"),
    "rhie_chow_density_dropped": (CLOSED, "", "mflow.faceE(i, j, k) *= rho.faceE(i, j, k);"),
    "nonorth_west_sign": (NONORTH, "b -= muW * dot(gradU.faceW(i, j, k), M.cW(i, j, k));",
                          "b += muW * dot(gradU.faceW(i, j, k), M.cW(i, j, k));"),
    "nonorth_c_swapped": (NONORTH, "b += muN * dot(gradU.faceN(i, j, k), M.cS(i, j, k));",
                          "b += muN * dot(gradU.faceN(i, j, k), M.cN(i, j, k));"),
    "nonorth_boundary_faces_unset": (NONORTH, "            if (false) gradU.faceW(i, j, k) = g;",
                                     "            if (i == 0) gradU.faceW(i, j, k) = g;"),
}
# the boundary fold b -= a_f * phi_patch (every FIXED face): its sign, and the term itself
FOLD_SIGN, FOLD_DROP = "fold_sign", "fold_dropped"


def _fold(src, drop):
    import re
    pat = re.compile(r"b -= a\[([1-6])\] \* F\.p(.)\(([^)]*)\)")
    out, n = pat.subn((lambda m: f"b -= 0. * F.p{m.group(2)}({m.group(3)})") if drop else
                      (lambda m: f"b += a[{m.group(1)}] * F.p{m.group(2)}({m.group(3)})"), src)
    return out, n


def _run_pins(lib, pins):
    env = dict(os.environ, ORACLE_LIB=lib, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "tests.oracle_lib_override", "-p", "no:cacheprovider",
                        pins], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-2000:]


def _build(src_text, tmp):
    cpp = os.path.join(tmp, "caffe_oracle.cpp")
    with open(cpp, "w") as f:
        f.write(src_text)
    shutil.copy(os.path.join(ROOT, "oracle", "caffe_oracle.h"), tmp)
    lib = os.path.join(tmp, "liboracle.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-shared", "-I", tmp, "-o", lib, cpp],
                   check=True, capture_output=True, timeout=300)
    return lib


@pytest.fixture(scope="module")
def source():
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    with open(SRC) as f:
        return f.read()


@pytest.mark.parametrize("pins", [CLOSED, NONORTH])
def test_clean_oracle_passes_the_pins(source, pins):
    with tempfile.TemporaryDirectory() as tmp:
        rc, out = _run_pins(_build(source, tmp), pins)
    assert rc == 0, out


@pytest.mark.parametrize("name", sorted(MUTANTS) + [FOLD_SIGN, FOLD_DROP])
def test_planted_mistake_fails_a_pin(source, name):
    if name in (FOLD_SIGN, FOLD_DROP):
        text, n = _fold(source, name == FOLD_DROP)
        pins = CLOSED
        assert n >= 6, "the fixed-face fold moved: update the mutant"
    else:
        pins, new, old = MUTANTS[name]
        assert source.count(old) >= 1, f"{name}: the planted line is no longer in the oracle"
        text = source.replace(old, new)
    with tempfile.TemporaryDirectory() as tmp:
        rc, out = _run_pins(_build(text, tmp), pins)
    print(f"{name}: {out.strip().splitlines()[-1]}")
    assert rc != 0, f"{name}: no pin failed\n{out}"
