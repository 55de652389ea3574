"""Singlet LR-CCSD (EOM-CCSD) excitation energies, restated for the tests
(test infrastructure only).

The reference's LR-CCSD (src/methods/coupledCluster/lrccsd/mod_lrccsd.f90:
176-326) iterates a trial vector r with the projections of
proj-lrccsd-spin-hf-1/2.f90 (the CCSD Jacobian acting on r in spin
orbitals), the Rayleigh quotient omega = <d, r>/<r, r> and the step
r -= step (d - omega r), and reports omega at convergence: an eigenvalue of
the CCSD Jacobian, i.e. an LR-CCSD excitation energy; spin = 0 keeps the
singlet combination r1(alpha) = r1(beta) (:298-314).

Here the singlet Jacobian is formed in the closed-shell amplitude space --
singles t1[i,a] and the pair-symmetric doubles t2[i,j,a,b] = t2[j,i,b,a],
one coordinate per pair (i a) <= (j b) -- as the central difference of the
biorthogonal closed-shell residual ``oracle.dressed.omega`` along every
coordinate, and diagonalised densely.  Pinned in the tests by the exact
two-electron case: there EOM-CCSD is exact and its n(n+1)/2 - 1 singlet
excitation energies are those of the two-electron singlet full CI
(``singlets_two_electrons``).
"""
import numpy as np

from oracle import dressed


def pair_list(o, v):
    ne = o * v
    return [(m, n) for n in range(ne) for m in range(n + 1)]


def jacobian(C, F, G, o, t1, t2, h=1e-5):
    """Dense d Omega / d t over (singles, pair-symmetric doubles)."""
    v = C.shape[0] - o
    P = pair_list(o, v)
    n1 = o * v
    dim = n1 + len(P)
    ii = np.array([m // v for m, n in P]); aa = np.array([m % v for m, n in P])
    jj = np.array([n // v for m, n in P]); bb = np.array([n % v for m, n in P])

    def vec(w1, w2):
        return np.concatenate([w1.ravel(), w2[ii, jj, aa, bb]])

    J = np.empty((dim, dim))
    for c in range(dim):
        d1 = np.zeros((o, v))
        d2 = np.zeros((o, o, v, v))
        if c < n1:
            d1.ravel()[c] = 1.0
        else:
            k = c - n1
            d2[ii[k], jj[k], aa[k], bb[k]] = 1.0
            d2[jj[k], ii[k], bb[k], aa[k]] = 1.0
        wp = dressed.omega(C, F, G, o, t1 + h * d1, t2 + h * d2)
        wm = dressed.omega(C, F, G, o, t1 - h * d1, t2 - h * d2)
        J[:, c] = (vec(*wp) - vec(*wm)) / (2 * h)
    return J


def excitation_energies(C, F, G, o, t1, t2, h=1e-5, vectors=False):
    """All eigenvalues of the singlet Jacobian, real ones sorted, and the
    count of complex ones; with ``vectors``, also the right eigenvectors of
    the real ones as (r1[i, a], r2[i, j, a, b]) with r2 pair-symmetric (a
    pair coordinate moves both t2[i,j,a,b] and t2[j,i,b,a])."""
    J = jacobian(C, F, G, o, t1, t2, h)
    if not vectors:
        lam = np.linalg.eigvals(J)
        real = np.abs(lam.imag) <= 1e-8 * np.abs(lam.real)
        return np.sort(lam.real[real]), int((~real).sum())
    lam, X = np.linalg.eig(J)
    real = np.abs(lam.imag) <= 1e-8 * np.abs(lam.real)
    order = np.argsort(lam.real[real], kind="stable")
    v = C.shape[0] - o
    n1 = o * v
    P = pair_list(o, v)
    out = []
    for x in X[:, real][:, order].real.T:
        r2 = np.zeros((o, o, v, v))
        for k, (m, n) in enumerate(P):
            r2[m // v, n // v, m % v, n % v] = r2[n // v, m // v, n % v, m % v] = x[n1 + k]
        out.append((x[:n1].reshape(o, v), r2))
    return lam.real[real][order], int((~real).sum()), out


def sigma(C, F, G, o, t1, t2, r1, r2, h=1e-4):
    """J r by the same central difference, for full (pair-symmetric) r2."""
    wp = dressed.omega(C, F, G, o, t1 + h * r1, t2 + h * r2)
    wm = dressed.omega(C, F, G, o, t1 - h * r1, t2 - h * r2)
    return (wp[0] - wm[0]) / (2 * h), (wp[1] - wm[1]) / (2 * h)


def singlets_two_electrons(C, core, G):
    """Excitation energies of the two-electron singlet full CI in the MO basis
    (spatially symmetric products (phi_p phi_q + phi_q phi_p) / norm, p <= q):
    <pq|H|rs> = h_pr d_qs + d_pr h_qs + (pr|qs) for the product phi_p(1) phi_q(2)."""
    n = C.shape[0]
    h = C.T @ core @ C
    X = np.einsum("mp,mq->mpq", C, C)
    g = np.einsum("mpq,mn,nrs->pqrs", X, G, X, optimize=True)
    basis = [(p, q) for p in range(n) for q in range(p, n)]

    def prod(p, q, r, s):
        return h[p, r] * (q == s) + (p == r) * h[q, s] + g[p, r, q, s]

    H = np.zeros((len(basis), len(basis)))
    for x, (p, q) in enumerate(basis):
        for y, (r, s) in enumerate(basis):
            nx = 1.0 / np.sqrt(2.0 * (1 + (p == q)))
            ny = 1.0 / np.sqrt(2.0 * (1 + (r == s)))
            H[x, y] = nx * ny * (prod(p, q, r, s) + prod(p, q, s, r) + prod(q, p, r, s) + prod(q, p, s, r))
    w = np.linalg.eigvalsh(H)
    return w[1:] - w[0]


def cis_spin_orbital(C, F, G, o):
    """Spin-orbital CIS matrix A[(i a), (j b)] = (e_a - e_i) d_ij d_ab + <aj||ib>
    over the 2o occupied and 2v virtual spin orbitals (alpha, beta alternating,
    as the reference numbers them), with <aj||ib> = R(a, i, j, b) of the
    reference's spin_hf_int (mod_lrccsd.f90:1449-1478).  Its eigenvalues are the
    singlet and the (threefold) triplet CIS energies together."""
    from oracle.triples import spin_integrals
    n = C.shape[0]
    R = spin_integrals(C, G)
    e = np.repeat(np.diag(F), 2)
    no, ns = 2 * o, 2 * n
    occ, vir = np.arange(no), np.arange(no, ns)
    A = R[np.ix_(vir, occ, occ, vir)].transpose(1, 0, 2, 3).reshape(no * (ns - no), no * (ns - no))
    A[np.diag_indices(A.shape[0])] += (e[vir][None, :] - e[occ][:, None]).ravel()
    return A


def average_excitation_level(r1, r2):
    """(S1 + 2 S2) / (S1 + S2) with S1 = sum r1^2, S2 = sum over i<j, a<b of r2^2
    (mod_lrccsd.f90:340-363), on the spin-orbital vector that the closed-shell
    (r1, r2) converts to (convertSpatialToSpinCC, mod_cc.f90:2601-2678)."""
    from oracle.triples import spin_amplitudes
    s1v, s2v = spin_amplitudes(r1, r2)
    no, nv = s1v.shape
    s1 = float(np.sum(s1v ** 2))
    s2 = 0.0
    for i in range(no - 1):
        for j in range(i + 1, no):
            for a in range(nv - 1):
                s2 += float(np.sum(s2v[i, j, a, a + 1:] ** 2))
    return (s1 + 2 * s2) / (s1 + s2)


def reference_weight(C, F, G, o, t1, omega, r1, r2):
    """r0 of energyLR (mod_lrccsd.f90:365-379) over spin orbitals: the vector
    converted as the amplitudes are (convertSpatialToSpinCC), normalised as
    ``normalize`` does (mod_lrccsd.f90:393-440: sum r1^2 + sum_{i<j,a<b} r2^2
    = 1); Fia(i,a) = F(i,a) + sum_kd t1(k,d) R(i,a,k,d)
    (proj-lrccsd-spin-hf-intermediates.f90:369-382); r0 = (sum_ia Fia r1 +
    sum_ijab R(j,b,i,a) r2(i,j,a,b) / 4) / omega with R = spin_hf_int.  Its
    magnitude (the eigenvector's sign is arbitrary)."""
    from oracle.triples import spin_amplitudes, spin_integrals
    n = C.shape[0]
    no, ns = 2 * o, 2 * n
    R = spin_integrals(C, G)
    Fs = np.zeros((ns, ns))
    Fs[0::2, 0::2] = F
    Fs[1::2, 1::2] = F
    st1, _ = spin_amplitudes(t1, np.zeros((o, o) + (n - o,) * 2))
    s1, s2 = spin_amplitudes(r1, r2)
    norm = np.sum(s1 ** 2)
    for i in range(no - 1):
        for j in range(i + 1, no):
            for a in range(ns - no - 1):
                norm += np.sum(s2[i, j, a, a + 1:] ** 2)
    scale = 1.0 / np.sqrt(norm)
    s1, s2 = s1 * scale, s2 * scale
    occ, vir = np.arange(no), np.arange(no, ns)
    Rovov = R[np.ix_(occ, vir, occ, vir)]
    Fia = Fs[np.ix_(occ, vir)] + np.einsum("kd,iakd->ia", st1, Rovov)
    sum1 = np.sum(Fia * s1)
    sum2 = np.einsum("jbia,ijab->", Rovov, s2)
    return abs((sum1 + sum2 / 4.0) / omega)
