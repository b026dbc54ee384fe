"""Case, mesh and solution files (caffe_old_amd.caseio) against the reference's file
formats: the INFO grammar of Boost's property tree as the reference's readers use it,
the example case decks (tests/golden/cases, copied from ExampleCases), the hexaMesh
grid writer and reader, and the Tecplot solution writer's exact bytes."""
import os
import shutil

import numpy as np
import pytest

from caffe_old_amd import caseio as io
from caffe_old_amd.case import tecplot_roundtrip, transfinite_nodes

CASES = os.path.join(os.path.dirname(__file__), "golden", "cases")


def _copy_case(src, dst):
    """Copy a deck with fresh, writable files and directories (the source tree may be
    read-only)."""
    for root, _, files in os.walk(src):
        rel = os.path.relpath(root, src)
        os.makedirs(os.path.join(dst, rel), exist_ok=True)
        for f in files:
            shutil.copyfile(os.path.join(root, f), os.path.join(dst, rel, f))


def _case(name):
    return os.path.join(CASES, name)


# ------------------------------------------------------------------ INFO grammar
def test_info_key_data_and_blocks():
    t = io.parse_info('; comment\nA 1 B 2\nC\n{\n  D "two words" ; trailing\n  E { F 3 }\n}\nG x\n')
    assert [c.key for c in t.children] == ["A", "B", "C", "G"]
    assert t.get("A", int) == 1 and t.get("B", int) == 2
    assert t.get("C.D") == "two words"
    assert t.get("C.E.F", float) == 3.0
    assert t.get("G") == "x"


def test_info_key_equals_value_reads_like_boost():
    # "MetricConversion = 0.1" (Cavity3D initialConditions.info): the data is "=" and
    # "0.1" becomes a key of its own, as Boost's parser alternates key and data
    t = io.parse_info("MetricConversion = 0.1\n")
    assert t.get("MetricConversion") == "="
    assert t.children[1].key == "0.1" and t.children[1].data == ""


def test_info_escapes_and_errors():
    t = io.parse_info('s "a\\tb\\"c"\n')
    assert t.get("s") == 'a\tb"c'
    with pytest.raises(io.InfoError):
        io.parse_info("A\n{\n B 1\n")
    with pytest.raises(io.InfoError):
        io.parse_info("}\n")
    with pytest.raises(io.InfoError):
        io.parse_info('s "open\n')


def test_info_conversions():
    t = io.parse_info("i 5\nf 1.\ng 2.5e-2\nh 5.0\nw abc\n")
    assert t.get("i", int) == 5
    assert t.get("f", float) == 1.0 and t.get("g", float) == 0.025
    with pytest.raises(io.InfoError):
        t.get("h", int)           # get<int> must consume the whole string
    with pytest.raises(io.InfoError):
        t.get("w", float)
    with pytest.raises(io.InfoError, match=r"No such node \(x.y\)"):
        t.get("x.y")


def test_info_include(tmp_path):
    (tmp_path / "inc.info").write_text("B 2\n")
    (tmp_path / "main.info").write_text('A 1\n#include "inc.info"\nC 3\n')
    t = io.read_info(str(tmp_path / "main.info"))
    assert [c.key for c in t.children] == ["A", "B", "C"]


def test_stov():
    assert io.stov("(1,0,0)") == (1.0, 0.0, 0.0)
    assert io.stov("(0, 0, 0)") == (0.0, 0.0, 0.0)
    assert io.stov("( 1.5 ,-2e-3, 3)") == (1.5, -0.002, 3.0)
    assert io.stov("(2,-0.6,1.6)") == (2.0, -0.6, 1.6)
    with pytest.raises(ValueError):
        io.stov("1,2,3")
    with pytest.raises(ValueError):
        io.stov("(1,2)")


# ------------------------------------------------------------------ the example decks
def test_cavity3d_deck():
    p = io.read_info(os.path.join(_case("Cavity3D"), "case", "case.info"))
    assert p.get("RunControl.maxNumberOfIterations", int) == 5
    assert p.get("RunControl.timeStep", float) == 2.5e-2
    assert p.get("Solver.timeAccurate") == "OFF"
    assert p.get("Boundaries.north.refVector") == "(10,0,0)"
    ics = io.read_info(os.path.join(_case("Cavity3D"), "case", "initialConditions.info"))
    nodes = io.generate_mesh(*io.read_mesh_info(os.path.join(_case("Cavity3D"), "mesh", "structuredMesh.info"))[1:])
    c = io.case_from_info(p, ics, nodes, "piso")
    assert not c.time_accurate and c.n_inner == 1 and c.n_pcorr == 2
    assert (c.alpha_u, c.alpha_p, c.rho, c.mu) == (0.5, 0.5, 1.0, 0.01)
    assert [b.type for b in c.boundaries] == ["wall"] * 6
    assert c.boundaries[2].refVector == (10.0, 0.0, 0.0) and c.boundaries[3].refVector == (-10.0, 0.0, 0.0)
    assert c.u_init == (0.0, 0.0, 0.0) and c.p_init == 0.0
    assert c.n_solves == 4                 # maxNumberOfIterations 5 runs 4 solves
    assert c.shape == (1680, 40, 40)
    s = io.case_from_info(p, ics, nodes, "simple")
    assert s.n_pcorr == 1


def test_unsteady_deck_and_sphere():
    p = io.read_info(os.path.join(_case("FlowOverIbSphere"), "case", "case.info"))
    ics = io.read_info(os.path.join(_case("FlowOverIbSphere"), "case", "initialConditions.info"))
    nodes = io.generate_mesh(*io.read_mesh_info(os.path.join(_case("FlowOverIbSphere"), "mesh", "structuredMesh.info"))[1:])
    c = io.case_from_info(p, ics, nodes, "ibpiso")
    assert c.time_accurate
    assert c.n_inner == p.get("Solver.numberOfInnerIterations", int)
    assert c.n_pcorr == p.get("Solver.numberOfPressureCorrections", int)
    assert c.u_init == (2.0, 0.0, 0.0)
    r, ctr = c.sphere
    assert r == p.get("ImmersedBoundaries.Sphere.radius", float)
    assert ctr == io.stov(p.get("ImmersedBoundaries.Sphere.center"))
    assert [b.type for b in c.boundaries][:2] == [p.get("Boundaries.east.type"), p.get("Boundaries.west.type")]


def test_decks_the_reference_rejects():
    # Channel3DWithIbSphere and SkewedCavity3D give u as refVector, so get<> throws on
    # the missing "u.value"; Cavity3DWithIbSphere's boundaries lack refValue, which the boundary
    # conditions read first (FlowBoundaryConditions.cpp:28).  This reader fails on the
    # same node.
    for name, missing in (("Channel3DWithIbSphere", "u.value"), ("SkewedCavity3D", "u.value"),
                          ("Cavity3DWithIbSphere", "Boundaries.east.refValue")):
        p = io.read_info(os.path.join(_case(name), "case", "case.info"))
        ics = io.read_info(os.path.join(_case(name), "case", "initialConditions.info"))
        nodes = io.generate_mesh(*io.read_mesh_info(os.path.join(_case(name), "mesh", "structuredMesh.info"))[1:])
        with pytest.raises(io.InfoError, match=missing.replace(".", r"\.")):
            io.case_from_info(p, ics, nodes, "ibpiso")


def test_every_deck_parses():
    for name in sorted(os.listdir(CASES)):
        d = _case(name)
        if not os.path.isdir(d):
            continue
        io.read_info(os.path.join(d, "case", "case.info"))
        io.read_info(os.path.join(d, "case", "initialConditions.info"))
        nm, v, res = io.read_mesh_info(os.path.join(d, "mesh", "structuredMesh.info"))
        assert nm and v.shape == (8, 3) and min(res) >= 2
        io.check_mesh(v)


# ------------------------------------------------------------------ meshes
def test_mesh_info_and_generation():
    name, v, res = io.read_mesh_info(os.path.join(_case("SkewedCavity3D"), "mesh", "structuredMesh.info"))
    assert res == (51, 31, 31)
    # MetricConversion 0.1 scales the vertices: v3 (1.8,1,0) -> (0.18, 0.1, 0)
    assert np.array_equal(v[2], 0.1 * np.array([1.8, 1.0, 0.0]))
    nd = io.generate_mesh(v, res)
    assert nd.shape == (31, 31, 51, 3)
    assert np.array_equal(nd, transfinite_nodes(v, 50, 30, 30))
    for (k, j, i), vi in (((0, 0, 0), 0), ((0, 0, 50), 1), ((0, 30, 50), 2), ((0, 30, 0), 3),
                          ((30, 0, 0), 4), ((30, 0, 50), 5), ((30, 30, 50), 6), ((30, 30, 0), 7)):
        assert np.allclose(nd[k, j, i], v[vi], rtol=0, atol=1e-15)
    # the west face is the plane x = 0 and the east face the line segment x = 0.1 + 0.08 y/0.1
    assert np.all(nd[:, :, 0, 0] == 0.0)
    assert np.allclose(nd[:, :, -1, 0], 0.1 + 0.8 * nd[:, :, -1, 1], atol=1e-15)


def test_check_mesh_rejects_a_twisted_face():
    v = np.array(io.read_mesh_info(os.path.join(_case("Cavity3D"), "mesh", "structuredMesh.info"))[1])
    io.check_mesh(v)
    v[6] += (0.0, 0.0, 1e-3)          # v7 off the east face's plane
    with pytest.raises(ValueError, match="not planar"):
        io.check_mesh(v)


def test_mesh_file_bytes_and_round_trip(tmp_path):
    nd = transfinite_nodes([(0, 0, 0), (1, 0, 0), (1, 1.0 / 3, 0), (0, 1.0 / 3, 0),
                            (0, 0, 2e-6), (1, 0, 2e-6), (1, 1.0 / 3, 2e-6), (0, 1.0 / 3, 2e-6)], 1, 1, 1)
    text = io.mesh_file_text("Box", nd)
    expected = ('TITLE = "Box"\nVARIABLES = "x", "y", "z"\nFILETYPE = GRID\nZONE I = 2, J = 2, K = 2\n'
                "DATAPACKING = BLOCK\n"
                "0 1 \n0 1 \n0 1 \n0 1 \n"
                "0 0 \n0.333333 0.333333 \n0 0 \n0.333333 0.333333 \n"
                "0 0 \n0 0 \n2e-06 2e-06 \n2e-06 2e-06 \n")
    assert text == expected
    f = tmp_path / "structuredMesh.dat"
    f.write_text(text)
    name, back = io.read_structured_mesh(str(f))
    assert name == "Box"
    assert np.array_equal(back, tecplot_roundtrip(nd))


def test_mesh_header_variants(tmp_path):
    f = tmp_path / "m.dat"
    f.write_text('title = "M"\n  anything at all\nZONE K = 2, I = 2, J = 2\nDATAPACKING = BLOCK\n' + "0 " * 24 + "\n")
    name, nd = io.read_structured_mesh(str(f))
    assert name == "M" and nd.shape == (2, 2, 2, 3)
    f.write_text('TITLE = "M"\nZONE I = 2, J = 2\nDATAPACKING = BLOCK\n')
    with pytest.raises(ValueError):
        io.read_structured_mesh(str(f))
    f.write_text('TITLE = "M"\nBOGUS 1\n')
    with pytest.raises(ValueError, match="invalid mesh header"):
        io.read_structured_mesh(str(f))


def test_hexamesh_module_writes_the_grid(tmp_path):
    from caffe_old_amd.run import hexamesh
    d = tmp_path / "SkewedCavity3D"
    _copy_case(_case("SkewedCavity3D"), str(d))
    out = hexamesh(str(d))
    name, nd = io.read_structured_mesh(out)
    _, v, res = io.read_mesh_info(str(d / "mesh" / "structuredMesh.info"))
    assert name == "SkewedCavity3D"
    assert np.array_equal(nd, tecplot_roundtrip(io.generate_mesh(v, res)))


# ------------------------------------------------------------------ run control
def test_run_control():
    t = io.parse_info("RunControl { terminationCondition iterations\n maxNumberOfIterations 5\n fileWriteInterval 2\n"
                      " timeStep 0.1\n maxSimTime 0.35 }\n")
    rc = io.RunControl(t)
    assert rc.n_solves() == 4
    writes = []
    while rc.continue_run():
        writes.append(rc.write_to_file())
    assert rc.itrs == 5 and writes == [False, True, False, True]
    t.get_child("RunControl.terminationCondition").data = "simTime"
    rc = io.RunControl(t)
    # 0.1, 0.2, 0.30000000000000004 < 0.35; 0.4 stops
    assert rc.n_solves() == 3
    t.get_child("RunControl.terminationCondition").data = "realTime"
    with pytest.raises(ValueError):
        io.RunControl(t)


# ------------------------------------------------------------------ solution file
def test_tec360_bytes():
    nd = transfinite_nodes([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], 1, 1, 1)
    w = io.Tec360Writer("unused", "Cube", nd, ["p", "massFlow"], ["u"])
    s1 = w.zone_text(0.025, {"p": np.array([[[1.0 / 3]]]), "massFlow": np.array([[[-0.0]]])},
                     {"u": [np.array([[[1e-20]]]), np.array([[[2.0]]]), np.array([[[-123456.789]]])]}, True)
    expected1 = ('TITLE = "Cube"\nVARIABLES = "x", "y", "z", "p", "massFlow", "u_x", "u_y", "u_z", '
                 'ZONE T = "CubeTime:0.025s"\nSTRANDID = 1, SOLUTIONTIME = 0.025\nI = 2, J = 2, K = 2\n'
                 "DATAPACKING = BLOCK\nVARLOCATION = ([4-8] = CELLCENTERED)\n"
                 "0 1 \n0 1 \n0 1 \n0 1 \n"
                 "0 0 \n1 1 \n0 0 \n1 1 \n"
                 "0 0 \n0 0 \n1 1 \n1 1 \n"
                 "0.3333333333333333 \n-0 \n9.999999999999999e-21 \n2 \n-123456.789 \n")
    assert s1 == expected1
    s2 = w.zone_text(0.05, {"p": np.array([[[0.5]]]), "massFlow": np.array([[[0.0]]])},
                     {"u": [np.array([[[0.0]]])] * 3}, False)
    expected2 = ('ZONE T = "CubeTime:0.05s"\nSTRANDID = 1, SOLUTIONTIME = 0.05\nI = 2, J = 2, K = 2\n'
                 "DATAPACKING = BLOCK\nVARSHARELIST = ([1-3] = 1)\nVARLOCATION = ([4-8] = CELLCENTERED)\n"
                 "0.5 \n0 \n0 \n0 \n0 \n")
    assert s2 == expected2
    one = io.Tec360Writer("unused", "C", nd, ["phi"], [])
    assert "VARLOCATION = ([4] = CELLCENTERED)\n" in one.zone_text(1.0, {"phi": np.zeros((1, 1, 1))}, {}, True)


def test_number_format_is_the_c_librarys():
    """The writers format doubles as an ostream does without fixed/scientific, i.e. C's
    %.*g; checked against the C library's own snprintf."""
    import ctypes
    libc = ctypes.CDLL(None)
    buf = ctypes.create_string_buffer(64)
    rng = np.random.default_rng(7)
    vals = list(rng.standard_normal(300) * 10.0 ** rng.integers(-30, 30, 300)) + \
        [0.0, -0.0, 1.0, 0.1, 1e-5, 1e5, 1e6, 123456.0, 1234567.0, 2.0 ** -1074, 1.7976931348623157e308]
    for prec in (6, 16):
        for v in vals:
            libc.snprintf(buf, 64, b"%.*g", ctypes.c_int(prec), ctypes.c_double(float(v)))
            assert io._g(float(v), prec) == buf.value.decode(), (prec, v)


def test_tec360_read_back(tmp_path):
    nd = transfinite_nodes([(0, 0, 0), (2, 0, 0), (2, 1, 0), (0, 1, 0), (0, 0, 1), (2, 0, 1), (2, 1, 1), (0, 1, 1)], 3, 2, 2)
    rng = np.random.default_rng(1)
    f = str(tmp_path / "s.dat")
    w = io.Tec360Writer(f, "S", nd, ["p", "massFlow"], ["u"])
    data = []
    for t in (0.5, 1.0):
        sc = {"p": rng.standard_normal((2, 2, 3)), "massFlow": rng.standard_normal((2, 2, 3))}
        ve = {"u": rng.standard_normal((3, 2, 2, 3))}
        w.write(t, sc, ve)
        data.append((t, sc, ve))
    nodes, zones = io.read_tec360(f)
    # the file holds 16 significant digits, which do not round-trip every double
    r16 = np.vectorize(lambda x: float("%.16g" % x))
    assert np.array_equal(nodes, r16(nd))
    assert len(zones) == 2
    for (t, sc, ve), z in zip(data, zones):
        assert z["time"] == t
        assert np.array_equal(z["p"], r16(sc["p"])) and np.array_equal(z["massFlow"], r16(sc["massFlow"]))
        for c, nm in enumerate(("u_x", "u_y", "u_z")):
            assert np.array_equal(z[nm], r16(ve["u"][c]))


def test_diffusion_deck():
    # FlatPlateDiffusion2D: the Diffusion solver's boundaries are fixed / empty with a
    # refValue each (PrimitiveBoundaryCondition.tpp:13-27), the initial field is phi
    d = _case("FlatPlateDiffusion2D")
    p = io.read_info(os.path.join(d, "case", "case.info"))
    ics = io.read_info(os.path.join(d, "case", "initialConditions.info"))
    name, v, res = io.read_mesh_info(os.path.join(d, "mesh", "structuredMesh.info"))
    c = io.case_from_info(p, ics, io.generate_mesh(v, (9, 5, 2)), "diffusion")
    assert c.algo == "diffusion" and not c.time_accurate
    assert [(b.type, b.refValue) for b in c.boundaries] == [("fixed", 0.0), ("fixed", 1.0), ("fixed", 0.0),
                                                            ("fixed", 0.0), ("empty", 0.0), ("empty", 0.0)]
    assert c.p_init == 0.0 and c.n_solves == 1
    # a flow solver fails on the deck where FlowBoundaryConditions does: it reads
    # refVector after refValue (FlowBoundaryConditions.cpp:27-29)
    with pytest.raises(io.InfoError, match=r"Boundaries\.east\.refVector"):
        io.case_from_info(p, ics, io.generate_mesh(v, (9, 5, 2)), "simple")


def test_coupled_deck_reads_only_its_keys():
    # The Coupled solver reads the flow boundaries, numberOfInnerIterations when
    # unsteady, relaxationFactorMomentum, rho and mu (Coupled.cpp:18-33); it reads no
    # relaxationFactorPCorr and no initial conditions
    d = _case("Cavity2D")
    p = io.read_info(os.path.join(d, "case", "case.info"))
    ics = io.read_info(os.path.join(d, "case", "initialConditions.info"))
    name, v, res = io.read_mesh_info(os.path.join(d, "mesh", "structuredMesh.info"))
    nodes = io.generate_mesh(v, (9, 9, 2))
    c = io.case_from_info(p, ics, nodes, "coupled")
    assert c.algo == "coupled" and c.time_accurate
    assert c.n_inner == p.get("Solver.numberOfInnerIterations", int)
    assert (c.alpha_u, c.rho, c.mu) == (0.5, 1.0, 0.001)
    assert c.u_init == (0.0, 0.0, 0.0) and c.p_init == 0.0
    assert [b.type for b in c.boundaries] == ["wall"] * 4 + ["empty"] * 2
    # a deck without relaxationFactorPCorr still loads for the Coupled solver, not for SIMPLE
    text = open(os.path.join(d, "case", "case.info")).read().replace("relaxationFactorPCorr 0.5", "")
    p2 = io.parse_info(text)
    io.case_from_info(p2, ics, nodes, "coupled")
    with pytest.raises(io.InfoError, match="relaxationFactorPCorr"):
        io.case_from_info(p2, ics, nodes, "simple")
