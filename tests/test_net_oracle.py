"""Pins of the FP64 oracle of the learned preconditioner (oracle/net_oracle.py).

Every linear piece is checked against its adjoint by the dot-product identity, a few
closed forms fix the conventions (tap order, channel order, interpolation weights), and
the full backward pass is checked against central finite differences of the loss.
"""
import numpy as np
import pytest

from oracle import net_oracle as no


def _rng(s):
    return np.random.default_rng(s)


def test_param_layout():
    assert no.param_count() == 27 * sum(ci * co for (_, ci, co, _, _) in no.LAYERS) == 512352
    th = np.arange(no.param_count(), dtype=float)
    W = no.split_params(th)
    assert W["in"].shape == (27, 16, 1) and W["out"].shape == (27, 1, 16)
    assert W["d2b"][0, 0, 0] == sum(27 * ci * co for (_, ci, co, _, _) in no.LAYERS[:5])


def test_conv_centre_tap_is_channel_product():
    x = _rng(0).standard_normal((4, 4, 4, 16))
    W = np.zeros((27, 32, 16))
    W[13] = _rng(1).standard_normal((32, 16))
    assert np.allclose(no.conv(W, x), x @ W[13].T)


def test_conv_tap_order():
    # tap t = (dz+1)*9 + (dy+1)*3 + (dx+1) reads x[k+dz][j+dy][i+dx]
    x = np.zeros((5, 5, 5, 1))
    x[2, 2, 3, 0] = 1.0          # one voxel east of the centre (i = 3)
    W = np.zeros((27, 1, 1))
    W[14, 0, 0] = 1.0            # dz=0, dy=0, dx=+1
    y = no.conv(W, x)
    assert y[2, 2, 2, 0] == 1.0 and abs(y).sum() == 1.0
    W = np.zeros((27, 1, 1))
    W[25, 0, 0] = 1.0            # dz=+1, dy=+1, dx=0 -> y[k][j][i] = x[k+1][j+1][i]
    x = np.zeros((5, 5, 5, 1))
    x[3, 3, 2, 0] = 1.0
    assert no.conv(W, x)[2, 2, 2, 0] == 1.0


def test_conv_zero_padding():
    x = np.ones((3, 3, 3, 1))
    W = np.ones((27, 1, 1))
    y = no.conv(W, x)[..., 0]
    assert y[1, 1, 1] == 27 and y[0, 0, 0] == 8 and y[0, 1, 1] == 18


@pytest.mark.parametrize("ci,co", [(1, 16), (16, 16), (32, 64), (16, 1)])
def test_conv_adjoints(ci, co):
    r = _rng(ci * 100 + co)
    x = r.standard_normal((4, 6, 8, ci))
    gy = r.standard_normal((4, 6, 8, co))
    W = r.standard_normal((27, co, ci))
    lhs = np.sum(no.conv(W, x) * gy)
    assert np.isclose(lhs, np.sum(x * no.conv_grad_input(W, gy)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(W * no.conv_grad_weight(x, gy)), rtol=1e-12)


def test_pool_up_adjoints_and_constants():
    r = _rng(3)
    a = r.standard_normal((4, 6, 8, 3))
    g = r.standard_normal((2, 3, 4, 3))
    assert np.isclose(np.sum(no.pool(a) * g), np.sum(a * no.pool_adj(g)), rtol=1e-12)
    c = r.standard_normal((2, 3, 4, 3))
    G = r.standard_normal((4, 6, 8, 3))
    assert np.isclose(np.sum(no.up(c) * G), np.sum(c * no.up_adj(G)), rtol=1e-12)
    assert np.allclose(no.pool(np.full((4, 4, 4, 2), 2.5)), 2.5)


def test_prolongation():
    r = _rng(4)
    c = r.standard_normal((3, 4, 5))
    f = r.standard_normal((6, 8, 10))
    assert np.isclose(np.sum(no.prolong(c) * f), np.sum(c * no.restrict(f)), rtol=1e-12)
    assert np.allclose(no.prolong(np.full((3, 4, 5), 1.75)), 1.75)
    # linear data in the interior is reproduced exactly: coarse centres at 2I+0.5
    I = np.arange(5, dtype=float)
    c = np.broadcast_to(2 * I + 0.5, (3, 4, 5)).copy()
    fx = no.prolong(c)[1, 2]
    assert np.allclose(fx[1:-1], np.arange(10)[1:-1])
    # 1-D weights: fine 2 -> 3/4 coarse 1 + 1/4 coarse 0 ; fine 0 clamps to coarse 0
    M = no._pweights(10, 5)
    assert M[2, 1] == 0.75 and M[2, 0] == 0.25 and M[0, 0] == 1.0 and M[9, 4] == 1.0


def test_spmv_matrix_properties():
    n = (8, 8, 8)
    A = no.random_pcorr_system(n, 1)
    one = no.pad(np.ones((8, 8, 8)), n)
    assert np.abs(no.spmv(A, one, n)).max() < 1e-13           # Neumann: constants in the null space
    r = _rng(5)
    x = no.pad(r.standard_normal((8, 8, 8)), n)
    y = no.pad(r.standard_normal((8, 8, 8)), n)
    assert np.isclose(x @ no.spmv(A, y, n), y @ no.spmv(A, x, n), rtol=1e-12)
    assert x @ no.spmv(A, x, n) > 0
    # against an explicitly assembled dense matrix
    xi = no.interior(x, n).ravel()
    cols = []
    for e in range(512):
        v = np.zeros(512)
        v[e] = 1.0
        cols.append(no.interior(no.spmv(A, no.pad(v.reshape(8, 8, 8), n), n), n).ravel())
    Ad = np.stack(cols, 1)
    assert np.allclose(Ad, Ad.T)
    assert np.allclose(Ad @ xi, no.interior(no.spmv(A, x, n), n).ravel())


def test_forward_is_linear():
    th = no.init_params(0)
    r = _rng(6)
    x1, x2 = r.standard_normal((8, 8, 8)), r.standard_normal((8, 8, 8))
    y = no.forward(th, 2.0 * x1 - 3.0 * x2)
    assert np.allclose(y, 2.0 * no.forward(th, x1) - 3.0 * no.forward(th, x2), rtol=1e-10, atol=1e-12)


def test_backward_adjoint_identity():
    # <gy, net(theta)(x)> is linear in each layer's weights: d/dW of it is the
    # backward pass; check <grad, theta_layer> = <gy, y> for the last layer (homogeneous of degree 1)
    th = no.init_params(1)
    r = _rng(7)
    x = r.standard_normal((8, 8, 8))
    gy = r.standard_normal((8, 8, 8))
    g = no.backward(th, x, gy)
    y = no.forward(th, x)
    sizes = no.layer_sizes()
    o = sum(sizes[:-1])
    assert np.isclose(g[o:] @ th[o:], np.sum(gy * y), rtol=1e-10)
    assert np.isclose(g[:sizes[0]] @ th[:sizes[0]], np.sum(gy * y), rtol=1e-10)


@pytest.mark.parametrize("post", [1, 3])
def test_two_grid_loss_gradient_finite_differences(post):
    """central differences pin dL/dtheta, the transposed post-smoothing sweeps included"""
    n = (16, 16, 16)
    A = no.random_pcorr_system(n, 2)
    r = _rng(8)
    xs = no.pad(r.standard_normal((16, 16, 16)), n)
    b = no.spmv(A, xs, n)
    th = no.init_params(3) * 0.5
    L, g = no.two_grid_loss(th, A, b, xs, n, grad=True, post=post)
    sizes = np.cumsum([0] + no.layer_sizes())
    for li in range(len(no.LAYERS)):
        idx = r.integers(sizes[li], sizes[li + 1], size=3)
        for q in idx:
            h = 1e-4 * max(1.0, abs(th[q]))
            tp, tm = th.copy(), th.copy()
            tp[q] += h
            tm[q] -= h
            fd = (no.two_grid_loss(tp, A, b, xs, n, post=post) - no.two_grid_loss(tm, A, b, xs, n, post=post)) / (2 * h)
            assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g).max()), (no.LAYERS[li][0], q, fd, g[q])


def test_two_grid_cycle_closed_forms():
    n = (16, 16, 16)
    A = no.random_pcorr_system(n, 4)
    r = _rng(9)
    b = no.pad(r.standard_normal((16, 16, 16)), n)
    # zero network: the cycle reduces to two damped-Jacobi sweeps from zero
    z = no.apply_precond(np.zeros(no.param_count()), A, b, n)
    inv = 1.0 / np.where(A.reshape(4, -1)[0] != 0, A.reshape(4, -1)[0], np.inf)
    z1 = no.OMEGA * inv * b
    assert np.allclose(z, z1 + no.OMEGA * inv * (b - no.spmv(A, z1, n)))
    # the loss of the exact solution as the cycle output is zero; of zero output it is one
    xs = no.pad(r.standard_normal((16, 16, 16)), n)
    bb = no.spmv(A, xs, n)
    L0 = no.two_grid_loss(np.zeros(no.param_count()), A, bb, xs, n)
    assert 0.0 < L0 < 1.0        # damped Jacobi reduces the energy error


def test_post_sweeps_converge_to_the_solution():
    """With the network switched off the cycle with `post` sweeps is damped Jacobi from
    z1 = w D^-1 b; on a nonsingular system it converges to A^-1 b as post grows (a wrong
    sign, weight or missing D^-1 in the sweep would not), and each sweep contracts the
    A-norm error."""
    n = (16, 16, 16)   # the smallest grid the network's four levels take
    A = no.random_pcorr_system(n, 12, neumann=False)
    N = 16 ** 3
    dense = np.zeros((N, N))
    for col in range(N):
        e = np.zeros(N)
        e[col] = 1.0
        dense[:, col] = no.interior(no.spmv(A, no.pad(e.reshape(16, 16, 16), n), n), n).ravel()
    b = _rng(13).standard_normal((16, 16, 16))
    x = np.linalg.solve(dense, b.ravel())
    zero = np.zeros(no.param_count())
    z = no.precond_parts(zero, A, no.pad(b, n), n, post=4000)["z"]
    assert np.allclose(no.interior(z, n).ravel(), x, rtol=1e-10, atol=1e-12)
    errs = []
    for post in (1, 2, 3, 4):
        e = no.interior(no.precond_parts(zero, A, no.pad(b, n), n, post=post)["z"], n).ravel() - x
        errs.append(float(e @ dense @ e))
    assert all(b_ < a_ for a_, b_ in zip(errs, errs[1:])), errs


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5e-3, 3.0e38])
    r = no.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0
    assert abs(r[4] / v[4] - 1) < 2 ** -8
    w = no.bf16(np.random.default_rng(0).standard_normal(1000))
    assert np.all(no.bf16(w) == w)     # idempotent


def test_forward_bf16_emulation_is_close():
    th = no.init_params(5)
    x = _rng(11).standard_normal((8, 8, 8))
    y = no.forward(th, x)
    yb = no.forward(th, x, emulate_bf16=True)
    assert 1e-5 < np.linalg.norm(yb - y) / np.linalg.norm(y) < 3e-2
