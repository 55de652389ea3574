"""Determinant full CI for tiny PPP molecules (test infrastructure only).

Plays the part of the reference's site-basis FCI (src/methods/mod_fci.f90:
101-185) as a pin: for two electrons CCSD is exact, so the CCSD oracle must
reproduce this energy.  Written from second quantisation, not from the
reference's code: H = E_nuc + sum_{pq,s} h_pq a+_ps a_qs
+ 1/2 sum_{pqrs,s,s'} (pq|rs) a+_ps a+_rs' a_ss' a_qs, in the MO basis.
"""
import itertools
import numpy as np


def _apply(det, ops):
    """Apply a sequence of (p, create?) operators right-to-left to a bit
    string determinant; return (sign, det) or (0, None)."""
    sign = 1
    for p, create in reversed(ops):
        occ = (det >> p) & 1
        if create == occ:
            return 0, None
        n_below = bin(det & ((1 << p) - 1)).count("1")
        if n_below & 1:
            sign = -sign
        det ^= (1 << p)
    return sign, det


def fci_energy(C, G, core, nocc):
    n = C.shape[1]
    h = C.T @ core @ C
    X = np.einsum("mp,mq->mpq", C, C)
    g = np.einsum("mpq,mn,nrs->pqrs", X, G, X, optimize=True)
    enuc = (G.sum() - np.trace(G)) / 2.0
    # spin orbital 2p+s
    alphas = [sum(1 << (2 * p) for p in c) for c in itertools.combinations(range(n), nocc)]
    betas = [sum(1 << (2 * p + 1) for p in c) for c in itertools.combinations(range(n), nocc)]
    dets = [a | b for a in alphas for b in betas]
    index = {d: k for k, d in enumerate(dets)}
    H = np.zeros((len(dets), len(dets)))
    for k, d in enumerate(dets):
        for p in range(n):
            for q in range(n):
                if h[p, q] == 0.0:
                    continue
                for s in range(2):
                    sg, d2 = _apply(d, [(2 * p + s, True), (2 * q + s, False)])
                    if sg:
                        H[index[d2], k] += sg * h[p, q]
        for p in range(n):
            for q in range(n):
                for r in range(n):
                    for s_ in range(n):
                        val = 0.5 * g[p, q, r, s_]
                        if abs(val) < 1e-300:
                            continue
                        for s1 in range(2):
                            for s2 in range(2):
                                sg, d2 = _apply(d, [(2 * p + s1, True), (2 * r + s2, True),
                                                    (2 * s_ + s2, False), (2 * q + s1, False)])
                                if sg:
                                    H[index[d2], k] += sg * val
    w = np.linalg.eigvalsh(H)
    return float(w[0] + enuc)
