"""GPU parity of the learned pressure-correction preconditioner against its FP64 oracle
(oracle/net_oracle.py), through the C ABI.

Tolerances: BF16 operands (8 significant bits) on MFMA with FP32 accumulation.
  * forward vs the oracle that rounds to BF16 at the same points: rel L2 <= 5e-3
    (left: FP32 accumulation order and rounding-boundary flips);
  * forward vs the exact FP64 oracle: rel L2 <= 3e-2;
  * parameter gradients (input gradients staged in BF16) vs FP64: rel L2 <= 5e-2 per layer;
  * the two-grid cycle with the network switched off or set to the identity: FP32
    arithmetic only (plus one BF16 rounding of the coarse input), rel L2 <= 1e-5 / 5e-3.
"""
import numpy as np
import pytest

from oracle import net_oracle as no

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - np.asarray(b, float)) / max(np.linalg.norm(np.asarray(b, float)), 1e-300))


@pytest.fixture(scope="module")
def net():
    from caffe_old_amd.learned import PressureNet
    n = PressureNet(seed=1)
    yield n
    n.close()


def _theta(seed, scale=1.0):
    return (no.init_params(seed) * scale).astype(np.float32)


def _identity_theta():
    """in: channel 0 <- input; d0, u0: channel 0 <- channel 0; out: output <- channel 0;
    every other layer zero, so net(v) = v (up to one BF16 rounding)."""
    th = np.zeros(no.param_count(), dtype=np.float32)
    W = no.split_params(th)   # views into th
    W["in"][13, 0, 0] = 1.0
    W["d0"][13, 0, 0] = 1.0
    W["u0"][13, 0, 0] = 1.0
    W["out"][13, 0, 0] = 1.0
    return th


def test_param_count(net):
    assert net.n_params == no.param_count()


def test_params_roundtrip(net):
    th = _theta(7)
    net.set_params(th)
    assert np.array_equal(net.get_params(), th)


# [K][J][I]: cubic, small anisotropic, and the coarse grids of the 128x32x32 and
# 256x64x64 channels (I >= 32 takes the 16-channel row kernels at level 0 and 1)
@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 16, 24), (16, 16, 64), (32, 32, 128)])
def test_forward_matches_oracle(net, shape):
    th = _theta(3)
    net.set_params(th)
    x = np.random.default_rng(4).standard_normal(shape).astype(np.float32)
    y = net.forward(x)
    t64 = th.astype(np.float64)
    yb = no.forward(t64, x.astype(np.float64), emulate_bf16=True)
    ye = no.forward(t64, x.astype(np.float64))
    print(f"forward {shape}: vs bf16-emulating oracle {_rel(y, yb):.2e}, vs fp64 {_rel(y, ye):.2e}")
    assert _rel(y, yb) <= 5e-3
    assert _rel(y, ye) <= 3e-2


def test_forward_identity_network(net):
    net.set_params(_identity_theta())
    x = np.random.default_rng(5).standard_normal((8, 8, 8)).astype(np.float32)
    y = net.forward(x)
    assert _rel(y, no.bf16(x)) <= 1e-6


def test_backward_matches_oracle(net):
    th = _theta(8)
    net.set_params(th)
    r = np.random.default_rng(9)
    x = r.standard_normal((16, 16, 16)).astype(np.float32)
    gy = r.standard_normal((16, 16, 16)).astype(np.float32)
    g = net.backward(x, gy)
    ge = no.backward(th.astype(np.float64), x.astype(np.float64), gy.astype(np.float64))
    off = np.cumsum([0] + no.layer_sizes())
    errs = {}
    for li, (name, *_rest) in enumerate(no.LAYERS):
        errs[name] = _rel(g[off[li]:off[li + 1]], ge[off[li]:off[li + 1]])
    print("backward rel L2 per layer:", {k: f"{v:.1e}" for k, v in errs.items()})
    assert max(errs.values()) <= 5e-2, errs


def _system(n, seed):
    A = no.random_pcorr_system(n, seed)
    return A.astype(np.float32).astype(np.float64)


def test_precond_network_off(net):
    """theta = 0: the cycle is two damped-Jacobi sweeps (FP32 arithmetic)."""
    n = (32, 16, 16)
    A = _system(n, 21)
    net.set_params(np.zeros(no.param_count(), dtype=np.float32))
    r = no.pad(np.random.default_rng(22).standard_normal(n[::-1]), n).astype(np.float32).astype(np.float64)
    z = net.precond(n, A, r)
    ze = no.apply_precond(np.zeros(no.param_count()), A, r, n)
    assert _rel(z, ze) <= 1e-5


def test_precond_identity_network(net):
    """net = identity: pins the restriction P^T and the trilinear prolongation P."""
    n = (32, 32, 16)
    A = _system(n, 23)
    th = _identity_theta()
    net.set_params(th)
    r = no.pad(np.random.default_rng(24).standard_normal(n[::-1]), n).astype(np.float32).astype(np.float64)
    z = net.precond(n, A, r)
    ze = no.apply_precond(th.astype(np.float64), A, r, n)
    print(f"identity-network cycle rel L2 {_rel(z, ze):.2e}")
    assert _rel(z, ze) <= 5e-3


@pytest.mark.parametrize("post", [1, 3])
def test_precond_matches_oracle(net, post):
    n = (32, 32, 32)
    A = _system(n, 25)
    th = _theta(26, 0.5)
    net.set_params(th)
    net.set_post_sweeps(post)
    try:
        r = no.pad(np.random.default_rng(27).standard_normal(n[::-1]), n).astype(np.float32).astype(np.float64)
        z = net.precond(n, A, r)
    finally:
        net.set_post_sweeps(1)
    ze = no.precond_parts(th.astype(np.float64), A, r, n, post=post)["z"]
    print(f"cycle ({post} post sweeps) rel L2 {_rel(z, ze):.2e}")
    assert _rel(z, ze) <= 3e-2


@pytest.mark.parametrize("post", [1, 3])
def test_loss_and_gradient_match_oracle(net, post):
    n = (32, 32, 32)
    A = _system(n, 31)
    rng = np.random.default_rng(32)
    xs = no.pad(rng.standard_normal(n[::-1]), n).astype(np.float32).astype(np.float64)
    b = no.spmv(A, xs, n).astype(np.float32).astype(np.float64)
    th = _theta(33, 0.5)
    net.set_params(th)
    net.set_post_sweeps(post)
    try:
        L, g = net.loss_grad(n, A, b, xs)
        L2 = net.loss_grad(n, A, b, xs, grad=False)
    finally:
        net.set_post_sweeps(1)
    Le, ge = no.two_grid_loss(th.astype(np.float64), A, b, xs, n, grad=True, post=post)
    off = np.cumsum([0] + no.layer_sizes())
    errs = {name: _rel(g[off[i]:off[i + 1]], ge[off[i]:off[i + 1]]) for i, (name, *_r) in enumerate(no.LAYERS)}
    print(f"loss gpu {L:.6f} oracle {Le:.6f}; grad rel L2 per layer", {k: f"{v:.1e}" for k, v in errs.items()})
    assert abs(L - Le) <= 1e-2 * abs(Le)
    assert max(errs.values()) <= 5e-2, errs
    # loss alone (no gradient) is the same number
    assert abs(L2 - L) <= 1e-6 * abs(L)
