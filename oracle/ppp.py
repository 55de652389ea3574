"""PPP model build, restated from the reference (test infrastructure only).

Follows src/input/readMoleculeInformation.f90 and the constants of
src/modules/mod_helios.f90:41-43.

Pinning: ``gamma`` is pinned by the hand value of a two-site Ohno gamma and by
the Hubbard limit (tests/test_oracle_ppp.py); ``core`` by its definition on a
two-site molecule.
"""
import numpy as np

# mod_helios.f90:41-43
BOHR_RADIUS = 0.52917721067      # Angstrom
HARTREE_ENERGY = 27.21138602     # eV
# readMoleculeInformation.f90:223  trConst = 1/(a0*Eh)
TR_CONST = 1.0 / (BOHR_RADIUS * HARTREE_ENERGY)


def distances(coords):
    """Site-site distances (readMoleculeInformation.f90:67-79)."""
    x = np.asarray(coords, dtype=np.float64)
    d = x[:, None, :] - x[None, :, :]
    return np.sqrt((d * d).sum(-1))


def gamma(coords, U, kind="ohno-klopman"):
    """Two-centre Coulomb integrals gamma_mn (readMoleculeInformation.f90:216-264).

    ohno-klopman:      med / sqrt(1 + (k*med*r)^2),  med = (U_m+U_n)/2
    mataga-nishimoto:  med / (1 + k*med*r)
    hubbard:           diagonal U only
    """
    r = distances(coords)
    U = np.asarray(U, dtype=np.float64)
    n = len(U)
    med = (U[:, None] + U[None, :]) / 2.0
    if kind == "ohno-klopman":
        return med / np.sqrt(1.0 + (TR_CONST * med * r) ** 2)
    if kind == "mataga-nishimoto":
        return med / (1.0 + TR_CONST * med * r)
    if kind == "hubbard":
        g = np.zeros((n, n))
        g[np.diag_indices(n)] = U
        return g
    raise ValueError(kind)


def connect(alpha, bonds, n):
    """Hueckel/connectivity matrix: alpha on the diagonal, beta on bonds
    (readMoleculeInformation.f90:108-117)."""
    h = np.zeros((n, n))
    h[np.diag_indices(n)] = alpha
    for (i, j, beta) in bonds:
        h[i, j] = beta
        h[j, i] = beta
    return h


def core(hconnect, G, nels):
    """Core Hamiltonian H_kk = alpha_k - sum_{l!=k} n_l gamma_kl
    (readMoleculeInformation.f90:140-159)."""
    H = hconnect.copy()
    nels = np.asarray(nels, dtype=np.float64)
    for k in range(len(nels)):
        val = 0.0
        for l in range(len(nels)):
            if k == l:
                continue
            val += nels[l] * G[k, l]
        H[k, k] -= val
    return H


def build(coords, bonds, alpha=0.0, U=11.13, nels=1, kind="ohno-klopman"):
    """Return (G, core, huckel) for a molecule of identical sites."""
    n = len(coords)
    al = np.full(n, alpha, dtype=np.float64) if np.isscalar(alpha) else np.asarray(alpha, float)
    Us = np.full(n, U, dtype=np.float64) if np.isscalar(U) else np.asarray(U, float)
    ne = np.full(n, nels, dtype=np.float64) if np.isscalar(nels) else np.asarray(nels, float)
    G = gamma(coords, Us, kind)
    hc = connect(al, bonds, n)
    return G, core(hc, G, ne), hc
