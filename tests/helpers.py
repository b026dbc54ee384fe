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
