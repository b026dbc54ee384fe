"""FP64 numpy restatement of the learned multilevel pressure-correction preconditioner.

TEST INFRASTRUCTURE ONLY: this module is the checker for the HIP network kernels
(caffe_old_amd/csrc/net.hip) and the learned FCG stage; the product path never
imports it.

The learned stage has no counterpart in the reference (its pressure solve is PETSc
BiCGStab + SOR, SparseMatrix.cpp:86-106).  Its definition lives here and in the HIP
sources; parity is pinned by
  * adjoint (dot-product) identities of every linear piece: conv vs its transposed
    conv, pool vs up-sampling, trilinear prolongation P vs restriction P^T;
  * central finite differences of the FP64 loss with respect to the parameters;
  * closed forms: a conv with a single centre tap is a channel matrix product, P of a
    constant is that constant, pooling of a constant is that constant.

Layouts (shared with the HIP code)
  coarse scalar fields   [K][J][I] (I fastest), float
  activations            [K][J][I][C] channels-last
  layer weights          [27][Cout][Cin], tap t = (dz+1)*9 + (dy+1)*3 + (dx+1)
  parameter vector       layers concatenated in LAYERS order
  fine fields            padded (nK+2)(nJ+2)(nI+2), one-cell halo of zeros
  p' matrix              symmetric storage A[4][P]: aP, and the coefficient of the
                         west / south / bottom face of each padded cell (the east
                         neighbour's coefficient is A[1][p+1], etc.)

Network (12 3x3x3 convolutions, zero padding, no bias, identity activation, so the
map input -> output is linear as a preconditioner must be):
  level 0 = coarse grid, levels 1..3 halve each dimension (2x2x2 average pooling);
  up-sampling is nearest (each child takes its parent's value).
"""
import numpy as np

OMEGA = 0.8  # damped-Jacobi weight of the pre- and post-smoothing sweeps

# (name, cin, cout, level, input mode)  mode: 'plain' | 'pool' | 'upadd'
LAYERS = [
    ("in", 1, 16, 0, "plain"),
    ("d0", 16, 16, 0, "plain"),
    ("d1a", 16, 32, 1, "pool"),
    ("d1b", 32, 32, 1, "plain"),
    ("d2a", 32, 64, 2, "pool"),
    ("d2b", 64, 64, 2, "plain"),
    ("d3a", 64, 64, 3, "pool"),
    ("d3b", 64, 64, 3, "plain"),
    ("u2", 64, 32, 2, "upadd"),
    ("u1", 32, 16, 1, "upadd"),
    ("u0", 16, 16, 0, "upadd"),
    ("out", 16, 1, 0, "plain"),
]
# source activations of each layer: (main input, skip input or None)
SOURCES = {
    "in": ("x", None), "d0": ("in", None), "d1a": ("d0", None), "d1b": ("d1a", None),
    "d2a": ("d1b", None), "d2b": ("d2a", None), "d3a": ("d2b", None), "d3b": ("d3a", None),
    "u2": ("d3b", "d2b"), "u1": ("u2", "d1b"), "u0": ("u1", "d0"), "out": ("u0", None),
}


def layer_sizes():
    return [27 * ci * co for (_, ci, co, _, _) in LAYERS]


def param_count():
    return int(sum(layer_sizes()))


def split_params(theta):
    out, o = {}, 0
    for (name, ci, co, _, _), n in zip(LAYERS, layer_sizes()):
        out[name] = theta[o:o + n].reshape(27, co, ci)
        o += n
    return out


def init_params(seed):
    """Reference initialisation (the HIP library draws its own; tests pass parameters in)."""
    rng = np.random.default_rng(seed)
    parts = []
    for (_, ci, co, _, _) in LAYERS:
        parts.append(rng.standard_normal(27 * co * ci) / np.sqrt(27.0 * ci))
    return np.concatenate(parts)


# ---------------------------------------------------------------- linear pieces
def _shift(x, dz, dy, dx):
    """y[k][j][i] = x[k+dz][j+dy][i+dx], zero outside (x is [K][J][I][C])."""
    K, J, I = x.shape[:3]
    y = np.zeros_like(x)
    zs, ze = max(0, -dz), min(K, K - dz)
    ys, ye = max(0, -dy), min(J, J - dy)
    xs, xe = max(0, -dx), min(I, I - dx)
    if zs < ze and ys < ye and xs < xe:
        y[zs:ze, ys:ye, xs:xe] = x[zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return y


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def conv(W, x):
    """3x3x3 convolution, zero padding: y[v][co] = sum_t sum_ci W[t][co][ci] x[v+off(t)][ci]."""
    K, J, I, ci = x.shape
    co = W.shape[1]
    y = np.zeros((K, J, I, co))
    for t, (dz, dy, dx) in enumerate(TAPS):
        y += _shift(x, dz, dy, dx) @ W[t].T
    return y


def conv_grad_input(W, gy):
    """Adjoint of conv in its input: gx[v][ci] = sum_t sum_co W[t][co][ci] gy[v-off(t)][co]."""
    gx = np.zeros(gy.shape[:3] + (W.shape[2],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        gx += _shift(gy, -dz, -dy, -dx) @ W[t]
    return gx


def conv_grad_weight(x, gy):
    """dW[t][co][ci] = sum_v gy[v][co] x[v+off(t)][ci]."""
    ci, co = x.shape[3], gy.shape[3]
    gW = np.zeros((27, co, ci))
    g2 = gy.reshape(-1, co)
    for t, (dz, dy, dx) in enumerate(TAPS):
        gW[t] = g2.T @ _shift(x, dz, dy, dx).reshape(-1, ci)
    return gW


def pool(x):
    K, J, I, C = x.shape
    return x.reshape(K // 2, 2, J // 2, 2, I // 2, 2, C).mean(axis=(1, 3, 5))


def pool_adj(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_adj(g):
    K, J, I, C = g.shape
    return g.reshape(K // 2, 2, J // 2, 2, I // 2, 2, C).sum(axis=(1, 3, 5))


# ---------------------------------------------------------------- network
def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(theta, x, keep=False, emulate_bf16=False):
    """x: coarse input [K][J][I] -> coarse output [K][J][I]; with keep, also the staged
    inputs of every layer (for the backward pass).

    emulate_bf16 rounds where the MFMA path does: the weights of the inner (MFMA)
    layers, every staged input of an inner layer and every layer output except the
    last (the first and last layers run in FP32 on their inputs and weights)."""
    Wd = split_params(theta)
    acts = {"x": x[..., None]}
    staged = {}
    rnd = bf16 if emulate_bf16 else (lambda a: a)
    for (name, ci, co, lev, mode) in LAYERS:
        src, skip = SOURCES[name]
        a = acts[src]
        if mode == "pool":
            s = pool(a)
        elif mode == "upadd":
            s = up(a) + acts[skip]
        else:
            s = a
        inner = name not in ("in", "out")
        if inner:
            s = rnd(s)
        staged[name] = s
        W = rnd(Wd[name]) if inner else Wd[name]
        y = conv(W, s)
        acts[name] = y if name == "out" else rnd(y)
    y = acts["out"][..., 0]
    return (y, acts, staged) if keep else y


def backward(theta, x, gy):
    """Gradient of <gy, forward(theta, x)> with respect to theta (flat, same layout)."""
    Wd = split_params(theta)
    _, acts, staged = forward(theta, x, keep=True)
    grads = {n: np.zeros_like(a) for n, a in acts.items()}
    grads["out"] = gy[..., None].copy()
    gW = {}
    for (name, ci, co, lev, mode) in reversed(LAYERS):
        g = grads[name]
        gW[name] = conv_grad_weight(staged[name], g)
        gs = conv_grad_input(Wd[name], g)
        src, skip = SOURCES[name]
        if mode == "pool":
            grads[src] += pool_adj(gs)
        elif mode == "upadd":
            grads[src] += up_adj(gs)
            grads[skip] += gs
        else:
            grads[src] += gs
    return np.concatenate([gW[n].ravel() for (n, _, _, _, _) in LAYERS])


# ---------------------------------------------------------------- fine-level operators
def _pweights(nf, nc):
    """1-D cell-centred linear prolongation: fine f takes 3/4 of coarse f>>1 and 1/4 of
    its other neighbour (clamped to the grid, i.e. constant extrapolation)."""
    M = np.zeros((nf, nc))
    for f in range(nf):
        c0 = f >> 1
        c1 = c0 + 1 if (f & 1) else c0 - 1
        c1 = min(max(c1, 0), nc - 1)
        M[f, c0] += 0.75
        M[f, c1] += 0.25
    return M


def prolong(c):
    """coarse [K][J][I] -> fine [2K][2J][2I] (trilinear, cell-centred)."""
    K, J, I = c.shape
    Mz, My, Mx = _pweights(2 * K, K), _pweights(2 * J, J), _pweights(2 * I, I)
    return np.einsum("zk,yj,xi,kji->zyx", Mz, My, Mx, c)


def restrict(f):
    """P^T: fine [2K][2J][2I] -> coarse."""
    K2, J2, I2 = f.shape
    Mz, My, Mx = _pweights(K2, K2 // 2), _pweights(J2, J2 // 2), _pweights(I2, I2 // 2)
    return np.einsum("zk,yj,xi,zyx->kji", Mz, My, Mx, f)


def interior(v, n):
    nI, nJ, nK = n
    return v.reshape(nK + 2, nJ + 2, nI + 2)[1:-1, 1:-1, 1:-1]


def pad(v, n):
    nI, nJ, nK = n
    out = np.zeros((nK + 2, nJ + 2, nI + 2))
    out[1:-1, 1:-1, 1:-1] = v
    return out.ravel()


def spmv(A, x, n):
    """y = A x for the symmetric 4-array storage on the padded grid (halo rows -> 0)."""
    nI, nJ, nK = n
    sJ, sK = nI + 2, (nI + 2) * (nJ + 2)
    P = sK * (nK + 2)
    A = A.reshape(4, P)
    aP, aI, aJ, aK = A
    y = np.zeros(P)
    sl = slice(sK + sJ + 1, P - sK - sJ - 1)
    idx = np.arange(P)[sl]
    y[idx] = (aP[idx] * x[idx] + aI[idx] * x[idx - 1] + aI[idx + 1] * x[idx + 1]
              + aJ[idx] * x[idx - sJ] + aJ[idx + sJ] * x[idx + sJ]
              + aK[idx] * x[idx - sK] + aK[idx + sK] * x[idx + sK])
    y = pad(interior(y, n), n)
    return y


def precond_parts(theta, A, r, n, post=1, emulate_bf16=False):
    """The learned two-grid cycle for residual r (padded fine):
         z1 = w D^-1 r;  v = P^T D^-1 (r - A z1);  c = net(v);
         y = z1 + P c;   z = y + w D^-1 (r - A y),
       the last step `post` times (each sweep from the previous z).
       Returns dict of the intermediates."""
    P = A.size // 4
    aP = A.reshape(4, P)[0]
    inv = np.where(aP != 0, 1.0 / np.where(aP != 0, aP, 1.0), 0.0)
    z1 = OMEGA * inv * r
    wv = inv * (r - spmv(A, z1, n))
    v = restrict(interior(wv, n))
    c = forward(theta, v, emulate_bf16=emulate_bf16)
    y = z1 + pad(prolong(c), n)
    z = y
    for _ in range(post):
        z = z + OMEGA * inv * (r - spmv(A, z, n))
    return {"z1": z1, "w": wv, "v": v, "c": c, "y": y, "z": z, "inv": inv}


def apply_precond(theta, A, r, n):
    return precond_parts(theta, A, r, n)["z"]


def two_grid_loss(theta, A, b, xs, n, grad=False, post=1):
    """Relative energy-norm error of one learned two-grid cycle (`post` post-smoothing
    sweeps) applied to b, whose exact solution is xs:  L = e^T A e / xs^T A xs,
    e = xs - M b.  With grad, also dL/dtheta."""
    pp = precond_parts(theta, A, b, n, post=post)
    e = xs - pp["z"]
    Ae = spmv(A, e, n)
    nrm = float(xs @ spmv(A, xs, n))
    L = float(e @ Ae) / nrm
    if not grad:
        return L
    g = -2.0 * Ae / nrm                                   # dL/dz
    h = g
    for _ in range(post):                                 # dL/dy = (I - w A D^-1)^post g
        h = h - OMEGA * spmv(A, pp["inv"] * h, n)
    gc = restrict(interior(h, n))                         # dL/dc = P^T h
    return L, backward(theta, pp["v"], gc)


def random_pcorr_system(n, seed, neumann=True):
    """A symmetric 7-point system on the padded grid: positive-magnitude (negative) face
    coefficients on interior faces; domain faces carry zero (neumann, singular with the
    constants as null space) or a coefficient that only enters aP (Dirichlet-like).
    aP = -(sum of the six face coefficients of the row)."""
    nI, nJ, nK = n
    rng = np.random.default_rng(seed)
    sJ, sK = nI + 2, (nI + 2) * (nJ + 2)
    P = sK * (nK + 2)
    F = np.zeros((3, nK + 2, nJ + 2, nI + 2))   # F[d][k][j][i]: low face of padded cell
    for d in range(3):
        # low faces of interior cells and the high domain face (first halo cell above)
        shape = [nK, nJ, nI]
        ext = list(shape)
        ext[2 - d] += 1
        vals = -(0.5 + rng.random(ext))
        if neumann:
            sl = [slice(None)] * 3
            sl[2 - d] = 0
            vals[tuple(sl)] = 0.0
            sl[2 - d] = -1
            vals[tuple(sl)] = 0.0
        F[d, 1:1 + ext[0], 1:1 + ext[1], 1:1 + ext[2]] = vals
    aP = np.zeros((nK + 2, nJ + 2, nI + 2))
    inner = (slice(1, nK + 1), slice(1, nJ + 1), slice(1, nI + 1))
    aP[inner] -= F[0][inner] + F[0][1:nK + 1, 1:nJ + 1, 2:nI + 2]
    aP[inner] -= F[1][inner] + F[1][1:nK + 1, 2:nJ + 2, 1:nI + 1]
    aP[inner] -= F[2][inner] + F[2][2:nK + 2, 1:nJ + 1, 1:nI + 1]
    A = np.zeros((4, P))
    A[0] = aP.ravel()
    for d in range(3):
        A[1 + d] = F[d].ravel()
    return A.ravel()
