"""The committed golden vectors (tests/golden/oracle_fixtures.npz, written by
tests/golden/make_fixtures.py from the FP64 oracle) against the oracle as it builds now:
every SIMPLE coefficient, right-hand side and field and every ghost row reproduces bit for
bit, so a later change to the oracle cannot move the vectors the GPU is checked against
without this test failing.  The tiny convolution is also pinned independently of the
network oracle: scipy's 3-D correlation with zero padding, one kernel per channel pair."""
import os

import numpy as np
import pytest

from tests.golden.make_fixtures import OUT, ghost_case, simple_case
from tests.helpers import STATE, oracle_momentum_band, oracle_pcorr_sym, oracle_rhs_padded, padded_of_rows

pytestmark = pytest.mark.usefixtures("oracle_built")


@pytest.fixture(scope="module")
def fx():
    if not os.path.exists(OUT):
        pytest.fail("tests/golden/oracle_fixtures.npz is missing: run tests/golden/make_fixtures.py")
    return np.load(OUT)


@pytest.mark.parametrize("n", [4, 8])
def test_simple_fixtures_reproduce(fx, n):
    from oracle.pyoracle import Oracle
    c = simple_case(n)
    o = Oracle(c, rtol=1e-14, maxit=20000)
    for it in range(1, 6):
        key = f"simple{n}_it{it}"
        if it in (1, 5):
            for k in STATE:
                assert np.array_equal(o.get(k), fx[f"{key}_before_{k}"]), (key, k)
        o.save_u0()
        o.compute_momentum(c.time_step)
        if it in (1, 5):
            band, _ = oracle_momentum_band(o)
            assert np.array_equal(band, fx[f"{key}_band"])
            for q in range(3):
                assert np.array_equal(oracle_rhs_padded(o, q, o.P), fx[f"{key}_mrhs"][q])
        o.compute_pcorr()
        if it in (1, 5):
            sym, _ = oracle_pcorr_sym(o)
            assert np.array_equal(sym, fx[f"{key}_pcsym"])
            assert np.array_equal(oracle_rhs_padded(o, 0, o.P), fx[f"{key}_pcrhs"])
        o.correct()
        if it in (1, 5):
            for k in ("u", "p", "mI", "mJ", "mK"):
                assert np.array_equal(o.get(k), fx[f"{key}_after_{k}"]), (key, k)


def test_ghost_row_fixtures_reproduce(fx):
    from oracle.pyoracle import Oracle
    o = Oracle(ghost_case())
    o.save_u0()
    o.compute_ib_field()
    o.compute_ib_coeffs()
    rid, st = o.imap()
    assert np.array_equal(st.astype(np.int8), fx["ib16_status"])
    rows, _, _, pad = padded_of_rows(o)
    ghosts = np.flatnonzero(st == 1)
    assert np.array_equal(pad[ghosts], fx["ib16_cells"])
    for which, name in ((0, "M"), (1, "P")):
        rp, cols, vals = o.matrix(which)
        for q, g in enumerate(ghosts):
            r = rid[g]
            m = rp[r + 1] - rp[r]
            assert np.array_equal(rows[cols[rp[r]:rp[r + 1]]], fx[f"ib16_cols{name}"][q, :m])
            assert np.array_equal(vals[rp[r]:rp[r + 1]], fx[f"ib16_vals{name}"][q, :m])


def test_tiny_conv_fixture_against_scipy(fx):
    from scipy.ndimage import correlate
    from oracle import net_oracle
    W, x, gy = fx["conv_W"], fx["conv_x"], fx["conv_gy"]
    assert np.array_equal(net_oracle.conv(W, x), fx["conv_y"])
    assert np.array_equal(net_oracle.conv_grad_input(W, gy), fx["conv_gx"])
    assert np.array_equal(net_oracle.conv_grad_weight(x, gy), fx["conv_gW"])
    co, ci = W.shape[1], W.shape[2]
    y = np.zeros(x.shape[:3] + (co,))
    for o_ in range(co):
        for i_ in range(ci):
            k = W[:, o_, i_].reshape(3, 3, 3)  # tap (dz+1)*9 + (dy+1)*3 + (dx+1)
            y[..., o_] += correlate(x[..., i_], k, mode="constant", cval=0.0)
    assert np.allclose(y, fx["conv_y"], rtol=0, atol=1e-12)
    # the input gradient is the adjoint: <conv(W, x), gy> = <x, conv_grad_input(W, gy)>
    assert np.isclose(np.vdot(fx["conv_y"], gy), np.vdot(x, fx["conv_gx"]), rtol=1e-12)
    assert np.isclose(np.vdot(fx["conv_y"], gy), np.vdot(W, fx["conv_gW"]), rtol=1e-12)


def test_network_fixture_reproduces(fx):
    from oracle import net_oracle
    theta = np.load(os.path.join(os.path.dirname(OUT), "..", "..", "caffe_old_amd", "weights",
                                 "pcorr_net_cavity128.npy")).astype(np.float64)
    assert np.array_equal(net_oracle.forward(theta, fx["net_x"]), fx["net_y"])
    g = net_oracle.backward(theta, fx["net_x"], fx["net_gout"])
    assert np.array_equal(g[fx["net_gidx"]], fx["net_gval"])
