"""Example case directories run unchanged through the file front end
(caffe_old_amd.caseio, caffe_old_amd.run) on the GPU, checked against the oracle.

Each test copies an example deck (tests/golden/cases) into a temporary directory,
lowers the resolution and the iteration count so that the oracle finishes in seconds,
runs ``hexamesh`` and the solver module, reads the Tecplot solution file back and
compares its last zone with the oracle run on the Case that the same files describe.
The file holds 16 significant digits; fields are compared at 1e-6 relative L2 (p
without its mean)."""
import os
import re
import shutil

import numpy as np
import pytest

from caffe_old_amd import caseio as io
from tests.helpers import interior

pytestmark = pytest.mark.gpu

CASES = os.path.join(os.path.dirname(__file__), "golden", "cases")


def _copy_case(src, dst):
    """Copy a deck with fresh, writable files and directories (the source tree may be
    read-only)."""
    for root, _, files in os.walk(src):
        rel = os.path.relpath(root, src)
        os.makedirs(os.path.join(dst, rel), exist_ok=True)
        for f in files:
            shutil.copyfile(os.path.join(root, f), os.path.join(dst, rel, f))


def _prepare(tmp_path, name, resolution, max_iters, extra=None, fix_ics=False):
    d = tmp_path / name
    _copy_case(os.path.join(CASES, name), str(d))
    if fix_ics:   # the deck names the initial values refVector / refValue; the reader wants "value"
        ic = d / "case" / "initialConditions.info"
        ic.write_text(re.sub(r"\b(refVector|refValue)\b", "value", ic.read_text()))
    mi = d / "mesh" / "structuredMesh.info"
    mi.write_text(re.sub(r"Resolution \([^)]*\)", f"Resolution ({resolution[0]},{resolution[1]},{resolution[2]})", mi.read_text()))
    ci = d / "case" / "case.info"
    text = re.sub(r"maxNumberOfIterations \d+", f"maxNumberOfIterations {max_iters}", ci.read_text())
    for k, v in (extra or {}).items():
        text = re.sub(rf"{k} \S+", f"{k} {v}", text)
    ci.write_text(text)
    return str(d)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _check_against_oracle(case_dir, algo):
    from oracle.pyoracle import Oracle
    from caffe_old_amd.run import hexamesh, run
    hexamesh(case_dir)
    s, rc, path = run(algo, case_dir, rtol=1e-12, quiet=True)
    sh = s.shape
    s.close()
    name, case, _ = io.load_case(case_dir, algo)
    nodes, zones = io.read_tec360(path)
    assert np.array_equal(nodes, case.nodes) or np.allclose(nodes, case.nodes, rtol=1e-15, atol=1e-15)
    # one zone per iteration (fileWriteInterval 1) and one after the loop
    assert len(zones) == case.n_solves + 1
    assert zones[-1]["time"] == pytest.approx(rc.sim_time, rel=1e-15)
    o = Oracle(case, rtol=1e-13)
    for _ in range(case.n_solves):
        o.solve()
    z = zones[-1]
    uo = interior(o.get("u"), sh)
    po = interior(o.get("p"), sh)
    errs = {}
    for c, nm in enumerate(("u_x", "u_y", "u_z")):
        if np.linalg.norm(uo[c]) > 0:
            errs[nm] = _rel(z[nm], uo[c])
    errs["p"] = _rel(z["p"] - z["p"].mean(), po - po.mean())
    print(f"{name} ({algo}, {sh}): file vs oracle {errs}")
    assert max(errs.values()) <= 1e-6, errs


def test_cavity3d_deck_piso(tmp_path):
    # Cavity3D: steady, lid speeds +-10 on north/south, alpha 0.5, 2 correctors
    d = _prepare(tmp_path, "Cavity3D", (13, 11, 9), 4, {"fileWriteInterval": 1})
    _check_against_oracle(d, "piso")


def test_skewed_cavity_deck_simple(tmp_path):
    # SkewedCavity3D: transfinite mesh of a skewed box (MetricConversion 0.1), unsteady
    d = _prepare(tmp_path, "SkewedCavity3D", (11, 9, 9), 4, {"fileWriteInterval": 1}, fix_ics=True)
    _check_against_oracle(d, "simple")


def test_channel3d_deck_simple(tmp_path):
    # Channel3D: inlet (west), outlet (east), walls; steady SIMPLE
    d = _prepare(tmp_path, "Channel3D", (13, 7, 7), 4, {"fileWriteInterval": 1})
    _check_against_oracle(d, "simple")
