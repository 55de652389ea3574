"""PPP Hamiltonian and closed-shell RHF: host-side preparation of the solver's
inputs (C, F_mo, gamma, e_ref) for synthetic molecules.

The model follows the reference's molecule reader
(src/input/readMoleculeInformation.f90: Ohno-Klopman gamma :216-264 with
e^2/(4 pi eps0) = a0*Eh from src/modules/mod_helios.f90:41-43, core
Hamiltonian :140-159, Hueckel matrix :108-117).  The RHF uses Roothaan steps
with Pulay (commutator) DIIS from the Hueckel density: the same converged
solution as the reference's damped-density iteration
(src/methods/mod_scf.f90:106-136), reached in fewer steps.  MO signs are fixed
so that the largest-|C| entry of each orbital is positive.
"""
import numpy as np

A0_EH = 0.52917721067 * 27.21138602   # eV*Angstrom


def model(coords, bonds, alpha=0.0, U=11.13, nels=1, kind="ohno-klopman"):
    x = np.asarray(coords, dtype=np.float64)
    n = len(x)
    r = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    Us = np.full(n, float(U)) if np.isscalar(U) else np.asarray(U, float)
    med = 0.5 * (Us[:, None] + Us[None, :])
    if kind == "ohno-klopman":
        G = med / np.sqrt(1.0 + (med * r / A0_EH) ** 2)
    elif kind == "mataga-nishimoto":
        G = med / (1.0 + med * r / A0_EH)
    elif kind == "hubbard":
        G = np.diag(Us)
    else:
        raise ValueError(kind)
    hk = np.zeros((n, n))
    hk[np.diag_indices(n)] = alpha
    for i, j, b in bonds:
        hk[i, j] = hk[j, i] = b
    ne = np.full(n, float(nels)) if np.isscalar(nels) else np.asarray(nels, float)
    core = hk.copy()
    off = G * ne[None, :]
    core[np.diag_indices(n)] -= off.sum(1) - np.diag(off)
    return G, core, hk


def _fock(core, G, D):
    F = core - G * D
    F[np.diag_indices_from(F)] += 2.0 * (G @ np.diag(D))
    return F


def _gauge(V):
    idx = np.argmax(np.abs(V), axis=0)
    s = np.sign(V[idx, np.arange(V.shape[1])])
    s[s == 0] = 1.0
    return V * s


def rhf(core, G, huckel, nocc, tol=1e-11, maxiter=300, ndiis=8, diis_start=1e-3, step=0.05):
    """Closed-shell PPP RHF from the Hueckel guess.

    Plain Roothaan steps on the commutator-damped density (the reference's
    scheme, src/methods/mod_scf.f90:106-136 and :446-494) until the commutator
    falls below diis_start, then Pulay DIIS on the AO Fock matrices.  DIIS
    from the start can lock onto an oscillation in long polyenes (N >= 100).
    """
    _, V = np.linalg.eigh(huckel)
    D = V[:, :nocc] @ V[:, :nocc].T
    Fs, Es = [], []
    for it in range(1, maxiter + 1):
        F = _fock(core, G, D)
        err = F @ D - D @ F
        emax = np.abs(err).max()
        if emax < tol:
            break
        if emax > diis_start and not Fs:
            R = 2.0 * D - np.eye(len(D))
            Fm = (F @ R - R @ F) @ R
            _, V = np.linalg.eigh(_fock(core, G, D - step * Fm))
            D = V[:, :nocc] @ V[:, :nocc].T
            continue
        Fs.append(F)
        Es.append(err)
        Fs, Es = Fs[-ndiis:], Es[-ndiis:]
        k = len(Fs)
        if k > 1:
            B = -np.ones((k + 1, k + 1))
            B[k, k] = 0.0
            for a in range(k):
                for b in range(a, k):
                    B[a, b] = B[b, a] = (Es[a] * Es[b]).sum()
            rhs = np.zeros(k + 1)
            rhs[k] = -1.0
            try:
                c = np.linalg.solve(B, rhs)[:k]
                F = sum(ci * Fi for ci, Fi in zip(c, Fs))
            except np.linalg.LinAlgError:
                pass
        _, V = np.linalg.eigh(F)
        D = V[:, :nocc] @ V[:, :nocc].T
    F = _fock(core, G, D)
    E, V = np.linalg.eigh(F)
    V = _gauge(V)
    D = V[:, :nocc] @ V[:, :nocc].T
    F = _fock(core, G, D)
    eref = float((D * (core + F)).sum() + (G.sum() - np.trace(G)) / 2.0)
    return dict(C=V, E=E, Fmo=V.T @ F @ V, eref=eref, iterations=it, error=float(np.abs(F @ D - D @ F).max()))


def prepare(mol, kind="ohno-klopman"):
    """Solver inputs of a molecules.Molecule: dict(C, F, G, nocc, eref, ...)."""
    G, core, hk = model(mol.coords, mol.bonds, mol.alpha, mol.U, 1, kind)
    r = rhf(core, G, hk, mol.nocc)
    return dict(C=r["C"], F=r["Fmo"], G=G, nocc=mol.nocc, eref=r["eref"], E=r["E"], core=core,
                scf_iterations=r["iterations"], scf_error=r["error"], name=mol.name)
