"""Closed-shell PPP RHF restated from the reference (test infrastructure only).

Follows src/methods/mod_scf.f90 (damped-density Roothaan iteration) and the
MO-Fock / reference-energy set-up of
src/methods/coupledCluster/mod_cc.f90:2311-2504.  The eigensolver is numpy's
``eigh`` with eigenvalues ascending, as the reference's tred4 sorts them
(src/modules/mod_math.f90:206-226); eigenvector signs are fixed so that the
largest-magnitude component of each column is positive.

Pinning: density idempotency, [F,D]=0, agreement of the converged energy with
a direct (undamped, tight) Roothaan solve, and alternant pairing of Hueckel
levels (tests/test_oracle_scf.py).
"""
import numpy as np


def fix_gauge(V):
    """Make the largest-|.| component of every column positive."""
    V = V.copy()
    idx = np.argmax(np.abs(V), axis=0)
    s = np.sign(V[idx, np.arange(V.shape[1])])
    s[s == 0] = 1.0
    return V * s


def eig(A):
    E, V = np.linalg.eigh(A)
    return E, fix_gauge(V)


def density(V, nocc):
    """D = C_occ C_occ^T (mod_scf.f90:390-413)."""
    Co = V[:, :nocc]
    return Co @ Co.T


def fockian(core, G, D):
    """F = core - G*D + 2 diag(G diag(D))  (mod_scf.f90:417-442)."""
    F = core - G * D
    F[np.diag_indices_from(F)] += 2.0 * (G @ np.diag(D))
    return F


def fmin(F, D):
    """Fmin = [F, 2D-1](2D-1)  (mod_scf.f90:446-494)."""
    R = 2.0 * D - np.eye(len(D))
    return (F @ R - R @ F) @ R


def energy(core, G, D, F):
    """E = sum D(core+F) + sum_{m!=n} G/2  (mod_scf.f90:166-200)."""
    eel = float((D * (core + F)).sum())
    enuc = float((G.sum() - np.trace(G)) / 2.0)
    return eel + enuc, eel, enuc


def rhf(core, G, huckel, nocc, accuracy=1e-12, maxiter=400, step=0.05):
    """Reference SCF: Hueckel guess, then damped-density Roothaan steps
    (mod_scf.f90:59-136, guess :351-386).

    Returns dict with C (ascending MO columns), E (orbital energies), D, F
    (AO Fock at convergence), energy, iterations, converged flag.
    """
    _, V = eig(huckel)
    D = density(V, nocc)
    F = fockian(core, G, D)
    it = 0
    E = None
    converged = False
    while True:
        it += 1
        Fm = fmin(F, D)
        eps = float(np.abs(Fm).max())
        D = D - step * Fm
        F = fockian(core, G, D)
        E, V = eig(F)
        D = density(V, nocc)
        F = fockian(core, G, D)
        if eps < accuracy:
            converged = True
            break
        if it + 1 > maxiter:
            break
    e_tot = energy(core, G, D, F)[0]
    return dict(C=V, E=E, D=D, F=F, energy=e_tot, iterations=it, converged=converged)


def mo_fock(C, core, G, nocc):
    """MO Fock matrix and reference energy as prepared for CC
    (mod_cc.f90:2311-2332 density, :2336-2396 Fock, :2474-2504 energy)."""
    D = density(C, nocc)
    X = fockian(core, G, D)
    eref = energy(core, G, D, X)[0]
    return C.T @ X @ C, eref
