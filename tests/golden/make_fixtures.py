"""Writes tests/golden/oracle_fixtures.npz: golden vectors emitted by the FP64 oracle.

Run from the repository root:  python tests/golden/make_fixtures.py

Contents (padded layout (nI+2)(nJ+2)(nK+2), i fastest; the oracle's CSR rows converted to
the device's band layouts by tests/helpers.py):
  * 4^3 and 8^3 SIMPLE lid cavities (Re 100, alpha_u 0.7, alpha_p 0.3; Simple.cpp:78-91):
    for iterations 1 and 5, the state before the iteration, the state after its momentum
    solve and Rhie-Chow, the momentum band and right-hand sides, the p' coefficients and
    right-hand side, and u, p, mI, mJ, mK after the iteration;
  * a 16^3 cavity with an immersed sphere: statuses and the 9-entry ghost rows of both
    matrices (IbPiso.cpp:65-190);
  * a tiny 3x3x3 convolution (3 -> 2 channels on 4^3) with its input and weight gradients,
    and the learned network (shipped 128^3 weights) on an 8^3 coarse grid: the forward
    output (FP64 and the BF16-emulating oracle) and, for a seeded output gradient, each
    layer's parameter-gradient norm and 64 sampled entries.
The oracle is test infrastructure: this script only records what it computes."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from caffe_old_amd import case as cs  # noqa: E402
from oracle import net_oracle  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402
from tests.helpers import STATE, oracle_momentum_band, oracle_pcorr_sym, oracle_rhs_padded  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_fixtures.npz")


def simple_case(n):
    return cs.lid_cavity(n, 0.01, 1.0, "simple", alpha_u=0.7, alpha_p=0.3, max_iterations=6)


def ghost_case():
    c = cs.lid_cavity(16, 0.01, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.02, n_pcorr=2,
                      max_iterations=2)
    c.algo = "ibpiso"
    c.sphere = (0.23, (0.47, 0.52, 0.5))
    return c


def main():
    f = {}
    for n in (4, 8):
        c = simple_case(n)
        o = Oracle(c, rtol=1e-14, maxit=20000)
        P = o.P
        for it in range(1, 6):
            before = {k: o.get(k) for k in STATE}
            o.save_u0()
            o.compute_momentum(c.time_step)
            band, _ = oracle_momentum_band(o)
            mrhs = np.stack([oracle_rhs_padded(o, q, P) for q in range(3)])
            mid = {k: o.get(k) for k in STATE}
            o.compute_pcorr()
            sym, _ = oracle_pcorr_sym(o)
            prhs = oracle_rhs_padded(o, 0, P)
            o.correct()
            if it in (1, 5):
                key = f"simple{n}_it{it}"
                for k, v in before.items():
                    f[f"{key}_before_{k}"] = v
                for k, v in mid.items():
                    f[f"{key}_mid_{k}"] = v
                f[f"{key}_band"] = band
                f[f"{key}_mrhs"] = mrhs
                f[f"{key}_pcsym"] = sym
                f[f"{key}_pcrhs"] = prhs
                for k in ("u", "p", "mI", "mJ", "mK"):
                    f[f"{key}_after_{k}"] = o.get(k)
    # immersed sphere ghost rows
    c = ghost_case()
    o = Oracle(c)
    o.save_u0()
    o.compute_ib_field()
    o.compute_ib_coeffs()
    rid, st = o.imap()
    f["ib16_status"] = st.astype(np.int8)
    from tests.helpers import padded_of_rows
    rows, _, _, pad = padded_of_rows(o)
    ghosts = np.flatnonzero(st == 1)
    cells = pad[ghosts]
    for which, name in ((0, "M"), (1, "P")):
        rp, cols, vals = o.matrix(which)
        C = np.full((len(ghosts), 9), -1, dtype=np.int64)
        V = np.zeros((len(ghosts), 9))
        for q, g in enumerate(ghosts):
            r = rid[g]
            cc = cols[rp[r]:rp[r + 1]]
            C[q, :len(cc)] = rows[cc]
            V[q, :len(cc)] = vals[rp[r]:rp[r + 1]]
        f[f"ib16_cols{name}"] = C
        f[f"ib16_vals{name}"] = V
    f["ib16_cells"] = cells.astype(np.int64)
    # a tiny convolution and its gradients
    rng = np.random.default_rng(11)
    W = rng.standard_normal((27, 2, 3))
    x = rng.standard_normal((4, 4, 4, 3))   # [K][J][I][C]
    gy = rng.standard_normal((4, 4, 4, 2))
    f["conv_W"], f["conv_x"], f["conv_gy"] = W, x, gy
    f["conv_y"] = net_oracle.conv(W, x)
    f["conv_gx"] = net_oracle.conv_grad_input(W, gy)
    f["conv_gW"] = net_oracle.conv_grad_weight(x, gy)
    # the network on an 8^3 coarse grid with the shipped weights
    theta = np.load(os.path.join(ROOT, "caffe_old_amd", "weights", "pcorr_net_cavity128.npy")).astype(np.float64)
    xin = rng.standard_normal((8, 8, 8))
    gout = rng.standard_normal((8, 8, 8))
    f["net_x"], f["net_gout"] = xin, gout
    f["net_y"] = net_oracle.forward(theta, xin)
    f["net_y_bf16"] = net_oracle.forward(theta, xin, emulate_bf16=True)
    g = net_oracle.backward(theta, xin, gout)
    sizes = net_oracle.layer_sizes()
    norms, idx, samp = [], [], []
    off = 0
    for sz in sizes:
        seg = g[off:off + sz]
        norms.append(np.linalg.norm(seg))
        pick = off + rng.choice(sz, size=min(64, sz), replace=False)
        idx.append(pick)
        samp.append(g[pick])
        off += sz
    f["net_gnorm"] = np.array(norms)
    f["net_gidx"] = np.concatenate(idx)
    f["net_gval"] = np.concatenate(samp)
    np.savez_compressed(OUT, **f)
    print(f"wrote {OUT}: {len(f)} arrays, {os.path.getsize(OUT) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
