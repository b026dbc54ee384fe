"""CPU tests that pin the FP64 oracle (oracle/caffe_oracle.cpp).

Each check is chosen so that a plausible restatement mistake (a dropped term, a wrong
sign, a transposed index, a wrong boundary distance) fails it:
  * the reference's own geometry code, compiled from its sources into oracle/_ref
    (Geometry.cpp, Sphere.cpp, Vector3D.cpp, Array3D.tpp): bit-identical results;
  * the reference's test known answers: GeometryTest.cpp:24-55, Array3DTest.cpp:54-92
    layout, IndexMapTest.cpp:36 row numbering, FieldTest.cpp:56-155 patch/face semantics;
  * closed forms: uniform-mesh metrics, exact gradients of linear fields, trilinear
    weights on boxes, the 7-point Laplacian;
  * invariants: p' matrix symmetric with zero row sums (pure Neumann); the momentum
    identity alpha_u a0 + sum a_nb - a0P = net outflow; post-correction imbalance = the
    linear residual b - A p'; dE(i) == dW(i+1) bit for bit;
  * brute force: dense LU solves of the assembled 4^3 systems.
"""
import numpy as np
import pytest

from caffe_old_amd import case as cs
from oracle import pyoracle as po
from tests.helpers import interior, oracle_momentum_band, oracle_pcorr_sym, padded_of_rows, strides

C = po.C


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _rand_hex(rng, skew=0.2):
    base = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], float)
    return np.ascontiguousarray(base * rng.uniform(0.5, 2.0, 3) + rng.uniform(-skew, skew, (8, 3)) + rng.normal(size=3))


@pytest.fixture(scope="module")
def ref():
    po.build()
    r = po.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return r


def test_geometry_bit_identical_to_reference_code(ref):
    L = po.lib()
    rng = np.random.default_rng(7)
    for _ in range(300):
        h = _rand_hex(rng)
        assert L.oc_hex_volume(_dp(h)) == ref.ref_hex_volume(_dp(h))
        a, b = np.zeros(3), np.zeros(3)
        L.oc_hex_centroid(_dp(h), _dp(a))
        ref.ref_hex_centroid(_dp(h), _dp(b))
        assert np.array_equal(a, b)
        q = np.ascontiguousarray(h[[0, 3, 7, 4]])
        assert L.oc_quad_area(_dp(q)) == ref.ref_quad_area(_dp(q))
        L.oc_quad_centroid(_dp(q), _dp(a))
        ref.ref_quad_centroid(_dp(q), _dp(b))
        assert np.array_equal(a, b)
        pt = np.ascontiguousarray(h.mean(0) + rng.uniform(-0.7, 0.7, 3))
        assert L.oc_inside_hex(_dp(pt), _dp(h)) == ref.ref_inside_hex(_dp(pt), _dp(h))


def test_mesh_metrics_match_reference_face_formula(ref):
    """HexaFvmMesh.cpp:725-729 face area normal, from the reference's own code, against the
    oracle's exported S arrays on a skewed mesh."""
    nodes = cs.perturb_interior(cs.cartesian_nodes(1, 1, 1, 4, 4, 4), 0.15, 0.25, seed=3)
    o = po.Oracle(cs.lid_cavity(4, 0.01, nodes=nodes))
    SI = o.metric("SI")
    _, sJ, sK, P = strides((4, 4, 4))
    for (i, j, k) in [(0, 0, 0), (2, 1, 3), (4, 3, 2), (1, 3, 3)]:
        q = np.ascontiguousarray(np.array([nodes[k, j, i], nodes[k, j + 1, i], nodes[k + 1, j + 1, i], nodes[k + 1, j, i]]))
        out = np.zeros(3)
        ref.ref_face_area_normal(_dp(q), _dp(out))
        p = (i + 1) + (j + 1) * sJ + (k + 1) * sK
        assert np.array_equal(SI[:, p], out)


def test_sphere_matches_reference(ref):
    L = po.lib()
    rng = np.random.default_rng(1)
    c = np.array([1.0, 0.5, 0.5])
    o = po.Oracle(cs.channel_ib_sphere(32, 8, 8, steps=1))
    del o
    for _ in range(200):
        pt = np.ascontiguousarray(c + rng.uniform(-0.2, 0.2, 3))
        a = np.zeros(3)
        ref.ref_sphere_nearest(0.1, _dp(c), _dp(pt), _dp(a))
        # Sphere::nearestIntersect = c + r unit(pt - c) (Sphere.cpp:23-30)
        d = pt - c
        inv = 1.0 / np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        expect = c + 0.1 * np.array([inv * d[0], inv * d[1], inv * d[2]])
        assert np.allclose(a, expect, rtol=0, atol=1e-15)
        inside = ref.ref_sphere_inside(0.1, _dp(c), _dp(pt))
        assert inside == (np.linalg.norm(pt - c) <= 0.1)
    del L


def test_geometry_known_answers():
    """GeometryTest.cpp:24-55: unit cube volume 1, unit quad area 1, (0.8,.5,.5) inside."""
    L = po.lib()
    cube = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], float)
    assert L.oc_hex_volume(_dp(cube)) == 1.0
    assert L.oc_quad_area(_dp(np.ascontiguousarray(cube[:4]))) == 1.0
    assert L.oc_inside_hex(_dp(np.array([0.8, 0.5, 0.5])), _dp(cube)) == 1
    assert L.oc_inside_hex(_dp(np.array([1.2, 0.5, 0.5])), _dp(cube)) == 0
    assert L.oc_inside_hex(_dp(np.array([1.0, 0.5, 0.5])), _dp(cube)) == 1  # faces count (Geometry.cpp:192)
    c = np.zeros(3)
    L.oc_hex_centroid(_dp(cube), _dp(c))
    assert np.allclose(c, 0.5, atol=1e-15)


def test_array3d_layout(ref):
    """Array3DTest.cpp:54-92: i fastest, then j, then k."""
    for (ni, nj, nk) in [(3, 4, 5), (7, 2, 3)]:
        q = 0
        for k in range(nk):
            for j in range(nj):
                for i in range(ni):
                    assert ref.ref_array3d_offset(ni, nj, nk, i, j, k) == q
                    q += 1


def test_index_map_row_numbering():
    """IndexMapTest.cpp:36 (serial): row = k*nJ*nI + j*nI + i."""
    o = po.Oracle(cs.lid_cavity(5, 0.01, nodes=cs.cartesian_nodes(1, 1, 1, 5, 4, 3)))
    rid, st = o.imap()
    assert np.array_equal(rid, np.arange(5 * 4 * 3))
    assert (st == 0).all()


def test_initial_patches_follow_change_type_order():
    """FieldTest.cpp:108-155 (FIXED patches = reference values) and the changeType order:
    setFixedBoundaries re-applies all refValues_ (PrimitiveBoundaryCondition.tpp:257-294)."""
    c = cs.lid_cavity(4, 0.01)
    c.boundaries = [cs.Boundary("wall", (1, 2, 3), 0), cs.Boundary("wall", (4, 5, 6), 0), cs.Boundary("wall", (7, 8, 9), 0),
                    cs.Boundary("wall", (1, 1, 1), 0), cs.Boundary("wall", (2, 2, 2), 0), cs.Boundary("wall", (3, 3, 3), 0)]
    o = po.Oracle(c)
    u = o.get("u").reshape(3, 6, 6, 6)
    assert (u[:, 2:4, 2:4, 5] == np.array([1, 2, 3])[:, None, None]).all()   # east
    assert (u[:, 2:4, 2:4, 0] == np.array([4, 5, 6])[:, None, None]).all()   # west
    assert (u[:, 2:4, 5, 2:4] == np.array([7, 8, 9])[:, None, None]).all()   # north
    assert (u[:, 0, 2:4, 2:4] == np.array([3, 3, 3])[:, None, None]).all()   # bottom
    # channel: outlet (east) first, then FIXED faces -> east patch of u = its refVector (0),
    # p east FIXED set before the other refValues exist
    ch = cs.channel_ib_sphere(8, 4, 4, steps=1)
    ch.boundaries[0] = cs.Boundary("outlet", (0, 0, 0), 2.5)
    o = po.Oracle(ch)
    p = o.get("p").reshape(6, 6, 10)
    assert (p[1:-1, 1:-1, -1] == 2.5).all() and (p[1:-1, 1:-1, 0] == 0).all()


def test_uniform_metrics_closed_form():
    n, h = 6, 1.0 / 6
    o = po.Oracle(cs.lid_cavity(n, 0.01))
    sh = (n, n, n)
    assert np.allclose(interior(o.metric("vol"), sh), h ** 3, rtol=1e-14, atol=0)
    dE, dW = interior(o.metric("dE"), sh), interior(o.metric("dW"), sh)
    assert np.allclose(dE[:, :, :-1], h, rtol=1e-13, atol=0) and np.allclose(dE[:, :, -1], 2 * h, rtol=1e-13, atol=0)
    assert np.allclose(dW[:, :, 1:], h, rtol=1e-13, atol=0) and np.allclose(dW[:, :, 0], 2 * h, rtol=1e-13, atol=0)
    gE = interior(o.metric("gE"), sh)
    assert np.allclose(gE[:, :, :-1], 0.5, rtol=1e-14) and (gE[:, :, -1] == 0).all()
    cE = interior(o.metric("cE"), sh)
    assert np.abs(cE).max() <= 1e-16


def test_shared_face_metrics_bitwise():
    """dE(i) == dW(i+1) and gW(i+1) + gE(i) = 1 on a skewed mesh."""
    nodes = cs.perturb_interior(cs.cartesian_nodes(1, 1, 1, 6, 5, 4), 0.15, 1 / 6, seed=5)
    o = po.Oracle(cs.lid_cavity(6, 0.01, nodes=nodes))
    sh = (6, 5, 4)
    for hi, lo, ax in (("dE", "dW", 2), ("dN", "dS", 1), ("dT", "dB", 0)):
        a, b = interior(o.metric(hi), sh), interior(o.metric(lo), sh)
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[ax] = slice(0, -1)
        sl_b[ax] = slice(1, None)
        assert np.array_equal(a[tuple(sl_a)], b[tuple(sl_b)])
    gE, gW = interior(o.metric("gE"), sh), interior(o.metric("gW"), sh)
    assert np.allclose(gE[:, :, :-1] + gW[:, :, 1:], 1.0, rtol=0, atol=5e-16)


def test_transfinite_generator_matches_oracle():
    """caffe_old_amd.case.transfinite_nodes (product) vs the oracle's restatement of
    HexaMeshGen::generateMesh: bit-identical."""
    v = np.array([(0, 0, 0), (1, 0, 0), (1.8, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1.8, 1, 1), (0, 1, 1)], float) * 0.1
    a = cs.transfinite_nodes(v, 7, 5, 4)
    b = np.zeros_like(a)
    po.lib().oc_transfinite_nodes(_dp(np.ascontiguousarray(v)), 8, 6, 5, _dp(b))
    assert np.array_equal(a, b)
    assert np.array_equal(cs.tecplot_roundtrip(np.array([0.9921875, 1 / 3])), np.array([0.992188, 0.333333]))


def test_trilinear_weights_closed_form():
    L = po.lib()
    pts = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], float) * 0.25 + 0.3
    rng = np.random.default_rng(2)
    for _ in range(20):
        xi = rng.uniform(0, 1, 3)
        x = np.ascontiguousarray(0.3 + 0.25 * xi)
        beta = np.zeros(8)
        L.oc_trilinear(_dp(np.ascontiguousarray(pts)), _dp(x), _dp(beta))
        a, b, c = xi
        expect = [(1 - a) * (1 - b) * (1 - c), a * (1 - b) * (1 - c), a * b * (1 - c), (1 - a) * b * (1 - c),
                  (1 - a) * (1 - b) * c, a * (1 - b) * c, a * b * c, (1 - a) * b * c]
        assert np.allclose(beta, expect, atol=1e-11)


def test_gradient_exact_for_linear_field():
    """FvScalarScheme divergence theorem: exact for linear fields at interior cells."""
    n = 6
    o = po.Oracle(cs.lid_cavity(n, 0.01))
    sh = (n, n, n)
    xc = o.metric("xc")
    lin = 0.3 * xc[0] - 1.7 * xc[1] + 2.2 * xc[2] + 4.0
    o.set("p", lin)
    o.set("pCorr", np.zeros_like(lin))
    o.set("gradPCorr", np.zeros((3, lin.size)))
    o.correct()  # p += 0; interpolates p and forms grad p
    g = interior(o.get("gradP"), sh)[:, 1:-1, 1:-1, 1:-1]
    assert np.allclose(g[0], 0.3, atol=1e-12) and np.allclose(g[1], -1.7, atol=1e-12) and np.allclose(g[2], 2.2, atol=1e-12)


def _state_after(case, iters):
    o = po.Oracle(case, rtol=1e-13)
    for _ in range(iters):
        o.solve()
    return o


def test_pcorr_matrix_symmetric_neumann_rows_sum_zero():
    o = _state_after(cs.lid_cavity(5, 0.01, max_iterations=4), 3)
    o.assemble_pcorr()
    sym, st = oracle_pcorr_sym(o)
    rp, cols, vals = o.matrix(1)
    n = len(rp) - 1
    A = np.zeros((n, n))
    for r in range(n):
        A[r, cols[rp[r]:rp[r + 1]]] = vals[rp[r]:rp[r + 1]]
    assert np.array_equal(A, A.T)
    assert np.abs(A.sum(1)).max() <= 1e-15 * np.abs(A).max()
    assert (np.diag(A) < 0).all()


def test_momentum_identity_net_outflow():
    """alpha_u a0 + sum a_nb - a0P = m_E - m_W + m_N - m_S + m_T - m_B for interior rows."""
    case = cs.lid_cavity(5, 0.01, algo="piso", alpha_u=0.7, alpha_p=0.3, time_accurate=True, dt=0.1, n_pcorr=2,
                         max_iterations=4)
    o = _state_after(case, 2)
    o.save_u0()
    o.assemble_momentum(case.time_step)
    band, st = oracle_momentum_band(o)
    sh = (5, 5, 5)
    _, sJ, sK, P = strides(sh)
    mI, mJ, mK = o.get("mI"), o.get("mJ"), o.get("mK")
    a0P = case.rho * (1 / 5) ** 3 / case.time_step
    pad = interior(np.arange(P), sh)[1:-1, 1:-1, 1:-1].ravel()
    lhs = 0.7 * band[0, pad] + band[1:, pad].sum(0) - a0P
    rhs = mI[pad + 1] - mI[pad] + mJ[pad + sJ] - mJ[pad] + mK[pad + sK] - mK[pad]
    assert np.allclose(lhs, rhs, rtol=0, atol=1e-14 * np.abs(band).max())


def test_post_correct_imbalance_is_linear_residual():
    """Simple.cpp:312-340: after correct(), the cell imbalance equals b - A p'."""
    case = cs.lid_cavity(5, 0.01, max_iterations=4)
    o = po.Oracle(case, ksp=po.KSP_PETSC_LIKE, rtol=1e-3)
    o.solve()
    o.save_u0()
    o.compute_momentum(case.time_step)
    o.compute_pcorr()
    rp, cols, vals = o.matrix(1)
    b = o.rhs(0)
    rows, rid, st, pad = padded_of_rows(o)
    x = o.get("pCorr")[rows]
    Ax = np.array([vals[rp[r]:rp[r + 1]] @ x[cols[rp[r]:rp[r + 1]]] for r in range(len(rows))])
    o.correct()
    m = o.get("massFlow")[rows]
    assert np.abs(m - (b - Ax)).max() <= 1e-12 * np.abs(b).max()
    assert np.abs(b - Ax).max() > 1e-6 * np.abs(b).max()  # loose solve: the residual is visible


def test_laplacian_special_case():
    """Steady, alpha = 1, zero flux: p' assembly is the 7-point Laplacian rho d h."""
    case = cs.lid_cavity(4, 0.01, alpha_u=1.0, alpha_p=1.0)
    o = po.Oracle(case)
    P = 6 ** 3
    o.set("d", np.full(P, 2.0))
    o.assemble_pcorr()
    sym, st = oracle_pcorr_sym(o)
    sh = (4, 4, 4)
    aI = interior(sym[1], sh)
    assert np.allclose(aI[:, :, 1:], 2.0 * 0.25, rtol=1e-14)
    assert (aI[:, :, 0] == 0).all()
    aP = interior(sym[0], sh)
    assert np.allclose(aP[1:-1, 1:-1, 1:-1], -6 * 0.5, rtol=1e-14)
    assert np.allclose(aP[0, 0, 0], -3 * 0.5, rtol=1e-14)  # corner: three Neumann faces folded


def test_dense_lu_brute_force():
    """The tight Krylov solution of the assembled 4^3 systems equals a dense LU solve."""
    case = cs.lid_cavity(4, 0.01, algo="piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.1, max_iterations=3)
    o = _state_after(case, 1)
    o.save_u0()
    o.compute_momentum(case.time_step)
    rp, cols, vals = o.matrix(0)
    n = len(rp) - 1
    A = np.zeros((n, n))
    for r in range(n):
        A[r, cols[rp[r]:rp[r + 1]]] = vals[rp[r]:rp[r + 1]]
    rows, rid, st, pad = padded_of_rows(o)
    ux = o.get("u")[0][rows]
    o.assemble_momentum(case.time_step)
    b = o.rhs(0)
    # b of the re-assembly differs (u* changed); solve the system the solve saw instead
    del b
    o2 = _state_after(case, 1)
    o2.save_u0()
    o2.assemble_momentum(case.time_step)
    b0 = o2.rhs(0)
    assert np.allclose(np.linalg.solve(A, b0), ux, rtol=0, atol=1e-11 * np.abs(ux).max())
    # pressure correction: singular (pure Neumann); compare modulo the constant
    o2.compute_momentum(case.time_step)
    o2.assemble_pcorr()
    rp, cols, vals = o2.matrix(1)
    Ap = np.zeros((n, n))
    for r in range(n):
        Ap[r, cols[rp[r]:rp[r + 1]]] = vals[rp[r]:rp[r + 1]]
    bp = o2.rhs(0)
    bp = bp - bp.mean()
    xp = np.linalg.lstsq(Ap, bp, rcond=None)[0]
    o2.compute_pcorr()
    xo = o2.get("pCorr")[rows]
    assert np.allclose(xo - xo.mean(), xp - xp.mean(), atol=1e-10 * np.abs(xp).max())


def test_continuity_scan_is_order_dependent():
    """Simple.cpp:370-373 compares |m| with the stored (normalised) value: a later larger
    imbalance does not replace an earlier one unless it exceeds |m_1|/V_1."""
    o = po.Oracle(cs.lid_cavity(4, 0.01))
    P = 6 ** 3
    m = np.zeros(P)
    sh = (4, 4, 4)
    pad = interior(np.arange(P), sh).ravel()
    V = 0.25 ** 3
    m[pad[0]] = 1e-6          # e = 1e-6 / V = 6.4e-5
    m[pad[5]] = 5e-6          # 5e-6 < 6.4e-5: skipped although its |m|/V is larger
    m[pad[9]] = -1e-4         # 1e-4 > 6.4e-5: e = 1e-4 / V
    m[pad[20]] = 2e-4         # 2e-4 < 6.4e-3: skipped
    o.set("massFlow", m)
    assert o.continuity() == pytest.approx(1e-4 / V, rel=1e-15)
    assert o.continuity_max() == pytest.approx(2e-4 / V, rel=1e-15)


def test_petsc_like_and_tight_agree_to_tolerance():
    case = cs.lid_cavity(6, 0.01, max_iterations=6)
    a = po.Oracle(case, ksp=po.KSP_TIGHT, rtol=1e-13)
    b = po.Oracle(case, ksp=po.KSP_PETSC_LIKE, rtol=1e-9, n_blocks=2, n_threads=2)
    for _ in range(5):
        a.solve()
        b.solve()
    sh = (6, 6, 6)
    ua, ub = interior(a.get("u"), sh), interior(b.get("u"), sh)
    assert np.linalg.norm(ua - ub) / np.linalg.norm(ua) < 1e-6


def test_ib_flags_and_ghost_rows():
    """IbPiso.cpp:65-190 on a channel: SOLID inside the sphere, ghosts have a FLUID
    neighbour, and on a box mesh the weights of each ghost row are trilinear (sum 1)."""
    case = cs.channel_ib_sphere(64, 16, 16, steps=1)
    case.sphere = (0.2, (1.0, 0.5, 0.5))
    o = po.Oracle(case)
    o.save_u0()
    o.compute_ib_field()
    o.compute_ib_coeffs()
    rid, st = o.imap()
    sh = (64, 16, 16)
    xc = interior(o.metric("xc"), sh).reshape(3, -1)
    r = np.linalg.norm(xc - np.array([1.0, 0.5, 0.5])[:, None], axis=0)
    assert ((st != 0) == (r <= 0.2)).all()
    st3 = st.reshape(16, 16, 64)
    ghost = np.argwhere(st3 == 1)
    for (k, j, i) in ghost:
        nb = [st3[k, j, i + 1], st3[k, j, i - 1], st3[k, j + 1, i], st3[k, j - 1, i], st3[k + 1, j, i], st3[k - 1, j, i]]
        assert 0 in nb
    rp, cols, vals = o.matrix(0)
    grows = rid[st == 1]
    for gr in grows:
        v = vals[rp[gr]:rp[gr + 1]]
        c = cols[rp[gr]:rp[gr + 1]]
        # (1 + beta_self) on the diagonal, beta_m elsewhere: the row sums to 2 when all
        # eight weights are kept (columns of SOLID cells are dropped)
        if len(c) == 9 or (len(c) == 8 and gr in c):
            assert abs(v.sum() - 2.0) < 1e-10


def test_ib_channel_relaxation_choice():
    """The IB channel's recorded relaxation (alpha_u 0.7, alpha_p 0.3; the configuration
    leaves it open): with 1 / 1 IbPiso (IbPiso.cpp:43-61, Piso's correctors without
    neighbour correction) diverges on the 64 x 16 x 16 channel within 15 steps, with
    0.7 / 0.3 the velocity stays bounded."""
    def max_u(au, ap, steps):
        c = cs.channel_ib_sphere(64, 16, 16, steps=steps)
        c.alpha_u, c.alpha_p = au, ap
        o = po.Oracle(c, rtol=1e-10, n_threads=4, n_blocks=4)
        for _ in range(steps):
            o.solve()
        return float(np.abs(interior(o.get("u"), c.shape)).max())
    assert max_u(1.0, 1.0, 15) > 100.0
    assert max_u(0.7, 0.3, 15) < 1.2
