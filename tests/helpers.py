"""Shared helpers for the parity tests: layout conversion between the oracle's
PETSc-style CSR rows (IndexMap row order) and the device's padded band layout."""
from __future__ import annotations

import numpy as np

from oracle.pyoracle import interior  # noqa: F401  (re-exported for tests)


def strides(shape):
    ni, nj, nk = shape
    sJ = ni + 2
    sK = (ni + 2) * (nj + 2)
    return 1, sJ, sK, (nk + 2) * sK


def padded_of_rows(o):
    """padded cell index of every matrix row of the oracle"""
    ni, nj, nk = o.shape
    rid, st = o.imap()
    _, sJ, sK, _ = strides(o.shape)
    q = np.arange(ni * nj * nk)
    i, j, k = q % ni, (q // ni) % nj, q // (ni * nj)
    pad = (i + 1) + (j + 1) * sJ + (k + 1) * sK
    rows = np.full(o.nrows(), -1, dtype=np.int64)
    live = rid >= 0
    rows[rid[live]] = pad[live]
    return rows, rid, st, pad


def oracle_momentum_band(o):
    """7 x P band (aP aE aW aN aS aT aB) of the oracle's A_[0]; ghost rows are left out
    (they are compared through the overlay)."""
    rp, cols, vals = o.matrix(0)
    rows, rid, st, pad = padded_of_rows(o)
    _, sJ, sK, P = strides(o.shape)
    off = {0: 0, 1: 1, -1: 2, sJ: 3, -sJ: 4, sK: 5, -sK: 6}
    band = np.zeros((7, P))
    status = np.full(P, 2)
    status[pad] = st
    for r in range(len(rows)):
        pr = rows[r]
        if status[pr] != 0:
            continue
        for e in range(rp[r], rp[r + 1]):
            dc = rows[cols[e]] - pr
            band[off[int(dc)], pr] = vals[e]
    return band, status


def oracle_pcorr_sym(o):
    """4 x P (aP, aI, aJ, aK) of the oracle's A_[1] for ACTIVE rows: aI at entry p is the
    coefficient between cells p-1 and p."""
    rp, cols, vals = o.matrix(1)
    rows, rid, st, pad = padded_of_rows(o)
    _, sJ, sK, P = strides(o.shape)
    out = np.zeros((4, P))
    status = np.full(P, 2)
    status[pad] = st
    for r in range(len(rows)):
        pr = rows[r]
        if status[pr] != 0:
            continue
        for e in range(rp[r], rp[r + 1]):
            dc = int(rows[cols[e]] - pr)
            if dc == 0:
                out[0, pr] = vals[e]
            elif dc == -1:
                out[1, pr] = vals[e]
            elif dc == 1:
                out[1, pr + 1] = vals[e]
            elif dc == -sJ:
                out[2, pr] = vals[e]
            elif dc == sJ:
                out[2, pr + sJ] = vals[e]
            elif dc == -sK:
                out[3, pr] = vals[e]
            elif dc == sK:
                out[3, pr + sK] = vals[e]
    return out, status


def oracle_rhs_padded(o, vec, P):
    rows, rid, st, pad = padded_of_rows(o)
    b = o.rhs(vec)
    out = np.zeros(P)
    out[rows] = b
    return out


def rel_inf(a, b, mask=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        a = a[..., mask]
        b = b[..., mask]
    den = np.abs(b).max()
    if den == 0:
        return float(np.abs(a - b).max())
    return float(np.abs(a - b).max() / den)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


STATE = ["u", "u0", "p", "pCorr", "d", "h", "gradP", "gradPCorr", "mI", "mJ", "mK", "massFlow"]


def copy_state(src, dst):
    """oracle -> device (or device -> oracle): every field the hot path reads"""
    for n in STATE:
        dst.set(n, src.get(n))


def threads():
    """host threads for the oracle: the CPU share this process may use, at most 16"""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def imbalance(b_padded, vol_padded, active):
    """max over active cells of |b| / V: the cell mass imbalance the p' assembly puts in
    its right-hand side (Simple.cpp:248-252), scaled as computeContinuityError scales it"""
    return float(np.max(np.abs(b_padded[active]) / vol_padded[active]))


def stepped_solve(g, o, dt, pc_opts, vol, active, n_pcorr):
    """Solver::solve(dt) of SIMPLE / PISO (Simple.cpp:78-91, Piso.cpp:12-29) run as its
    sub-steps on the device g and the oracle o side by side.  Returns the pre-correction
    imbalance of the first p' assembly on each side (the outer-residual history the
    physics determines: it does not depend on the Krylov path of the p' solve)."""
    g.save_u0()
    o.save_u0()
    g.momentum_assemble(dt)
    g.momentum_solve()
    g.rhie_chow()
    o.compute_momentum(dt)
    hg = ho = None
    for q in range(n_pcorr):
        g.pcorr_assemble()
        o.assemble_pcorr()
        if q == 0:
            hg = imbalance(g.rhs(1), vol, active)
            ho = imbalance(oracle_rhs_padded(o, 0, g.P), vol, active)
        g.pcorr_solve(**pc_opts)
        g.correct()
        o.compute_pcorr()
        o.correct()
    return hg, ho


def assembly_errors(g2, o, dt):
    """Momentum and p' assembly of the device context g2 and the oracle from the oracle's
    current state (copied into g2).  Leaves the oracle's state as it was.  Returns the
    max-norm relative differences of the band, the three momentum right-hand sides, d,
    the p' coefficients and the p' right-hand side."""
    snap = {n: o.get(n) for n in STATE}
    copy_state(o, g2)
    g2.momentum_assemble(dt)
    o.assemble_momentum(dt)
    band_o, status = oracle_momentum_band(o)
    act = status == 0
    out = {"band": rel_inf(g2.matrix(0), band_o, act)}
    bg = g2.rhs(0)
    for c in range(3):
        out[f"rhs{c}"] = rel_inf(bg[c], oracle_rhs_padded(o, c, g2.P), act)
    out["d"] = rel_inf(g2.get("d"), o.get("d"), act)
    for n, v in snap.items():
        o.set(n, v)
    copy_state(o, g2)
    g2.pcorr_assemble()
    o.assemble_pcorr()
    sym_o, _ = oracle_pcorr_sym(o)
    sym_g = g2.matrix(1)
    out["aP"] = rel_inf(sym_g[0], sym_o[0], act)
    out["aFaces"] = max(rel_inf(sym_g[a], sym_o[a]) for a in (1, 2, 3))
    out["pc_rhs"] = rel_inf(g2.rhs(1), oracle_rhs_padded(o, 0, g2.P), act)
    for n, v in snap.items():
        o.set(n, v)
    return out


def field_errors(g, o):
    """relative L2 of u, v, w over the interior cells, and of p after removing the mean
    (the pure-Neumann pressure is defined up to a constant)"""
    sh = g.shape
    ug, uo = interior(g.get("u"), sh), interior(o.get("u"), sh)
    pg, po = interior(g.get("p"), sh), interior(o.get("p"), sh)
    errs = {c: rel_l2(ug[i], uo[i]) for i, c in enumerate("uvw") if np.linalg.norm(uo[i]) > 0}
    errs["p"] = rel_l2(pg - pg.mean(), po - po.mean())
    return errs


def active_mask(o):
    """padded boolean mask of the ACTIVE cells of the oracle's IndexMap"""
    rows, rid, st, pad = padded_of_rows(o)
    m = np.zeros(o.P, dtype=bool)
    m[pad[st == 0]] = True
    return m
