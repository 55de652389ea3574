"""Independent spin-orbital CCSD (test infrastructure only).

Textbook equations of J. F. Stanton, J. Gauss, J. D. Watts, R. J. Bartlett,
J. Chem. Phys. 94, 4334 (1991), in antisymmetrised-integral form, with a
plain Jacobi update and no DIIS.  Shares nothing with ``oracle.ccsd`` except
the ZDO integral definition; it exists to pin the spatial-orbital oracle
(which restates the reference's hand-derived contravariant equations).

Spin orbitals are ordered (p, alpha), (p, beta) for spatial p; occupied
first.
"""
import numpy as np

einsum = lambda *a: np.einsum(*a, optimize=True)   # noqa: E731


def spin_integrals(C, G, Fmo, nocc):
    """<pq||rs> in physicists' notation and the spin Fock matrix."""
    n = C.shape[1]
    X = einsum("mp,mq->mpq", C, C)
    g = einsum("mpq,mn,nrs->pqrs", X, G, X)         # (pq|rs) chemists'
    # spin-orbital index P = 2p + s; reorder so occupied spin orbitals come first
    order = []
    for p in range(nocc):
        order += [2 * p, 2 * p + 1]
    for p in range(nocc, n):
        order += [2 * p, 2 * p + 1]
    order = np.array(order)
    sp = order // 2
    ss = order % 2
    ns = 2 * n
    # <PQ|RS> = (PR|QS) delta(sP,sR) delta(sQ,sS)
    phys = g[sp[:, None, None, None], sp[None, None, :, None], sp[None, :, None, None], sp[None, None, None, :]]
    # phys[P,Q,R,S] = (sp[P] sp[R] | sp[Q] sp[S])
    phys = phys * ((ss[:, None, None, None] == ss[None, None, :, None])
                   & (ss[None, :, None, None] == ss[None, None, None, :]))
    anti = phys - phys.transpose(0, 1, 3, 2)
    fs = Fmo[sp[:, None], sp[None, :]] * (ss[:, None] == ss[None, :])
    return anti, fs, 2 * nocc, ns


def ccsd(C, G, Fmo, nocc, eref, eps=1e-11, maxiter=500, t1=None, t2=None):
    W, f, no, ns = spin_integrals(C, G, Fmo, nocc)
    o, v = slice(0, no), slice(no, ns)
    nv = ns - no
    fo = np.diag(f)[o]
    fv = np.diag(f)[v]
    Dia = fo[:, None] - fv[None, :]
    Dijab = fo[:, None, None, None] + fo[None, :, None, None] - fv[None, None, :, None] - fv[None, None, None, :]
    if t1 is None:
        t1 = f[o, v] / Dia
    if t2 is None:
        t2 = W[o, o, v, v] / Dijab
    fd = f - np.diag(np.diag(f))
    it = 0
    for it in range(1, maxiter + 1):
        t1n, t2n = update(W, f, fd, t1, t2, o, v, no, nv)
        t1n /= Dia
        t2n /= Dijab
        diff = max(np.abs(t1n - t1).max(), np.abs(t2n - t2).max())
        t1, t2 = t1n, t2n
        if diff < eps:
            break
    e = energy(W, f, t1, t2, o, v)
    return e + eref, e, t1, t2, it


def energy(W, f, t1, t2, o, v):
    return float(einsum("ia,ia->", f[o, v], t1) + 0.25 * einsum("ijab,ijab->", W[o, o, v, v], t2)
                 + 0.5 * einsum("ijab,ia,jb->", W[o, o, v, v], t1, t1))


def update(W, f, fd, t1, t2, o, v, no, nv):
    """Numerators of the Stanton-Gauss T1/T2 equations (off-diagonal Fock)."""
    tau_t = t2 + 0.5 * (einsum("ia,jb->ijab", t1, t1) - einsum("ib,ja->ijab", t1, t1))
    tau = t2 + einsum("ia,jb->ijab", t1, t1) - einsum("ib,ja->ijab", t1, t1)
    Fae = fd[v, v] - 0.5 * einsum("me,ma->ae", f[o, v], t1) + einsum("mf,mafe->ae", t1, W[o, v, v, v]) \
        - 0.5 * einsum("mnaf,mnef->ae", tau_t, W[o, o, v, v])
    Fmi = fd[o, o] + 0.5 * einsum("me,ie->mi", f[o, v], t1) + einsum("ne,mnie->mi", t1, W[o, o, o, v]) \
        + 0.5 * einsum("inef,mnef->mi", tau_t, W[o, o, v, v])
    Fme = f[o, v] + einsum("nf,mnef->me", t1, W[o, o, v, v])
    Wmnij = W[o, o, o, o] + einsum("je,mnie->mnij", t1, W[o, o, o, v]) \
        - einsum("ie,mnje->mnij", t1, W[o, o, o, v]) + 0.25 * einsum("ijef,mnef->mnij", tau, W[o, o, v, v])
    Wabef = W[v, v, v, v] - einsum("mb,amef->abef", t1, W[v, o, v, v]) \
        + einsum("ma,bmef->abef", t1, W[v, o, v, v]) + 0.25 * einsum("mnab,mnef->abef", tau, W[o, o, v, v])
    Wmbej = W[o, v, v, o] + einsum("jf,mbef->mbej", t1, W[o, v, v, v]) \
        - einsum("nb,mnej->mbej", t1, W[o, o, v, o]) \
        - einsum("jnfb,mnef->mbej", 0.5 * t2 + einsum("jf,nb->jnfb", t1, t1), W[o, o, v, v])
    t1n = f[o, v] + einsum("ie,ae->ia", t1, Fae) - einsum("ma,mi->ia", t1, Fmi) \
        + einsum("imae,me->ia", t2, Fme) - einsum("nf,naif->ia", t1, W[o, v, o, v]) \
        - 0.5 * einsum("imef,maef->ia", t2, W[o, v, v, v]) - 0.5 * einsum("mnae,nmei->ia", t2, W[o, o, v, o])
    Fae2 = Fae - 0.5 * einsum("mb,me->be", t1, Fme)
    Fmi2 = Fmi + 0.5 * einsum("je,me->mj", t1, Fme)
    P = lambda x, a, b: x - x.swapaxes(a, b)                     # noqa: E731
    t2n = W[o, o, v, v].copy()
    t2n += P(einsum("ijae,be->ijab", t2, Fae2), 2, 3)
    t2n -= P(einsum("imab,mj->ijab", t2, Fmi2), 0, 1)
    t2n += 0.5 * einsum("mnab,mnij->ijab", tau, Wmnij)
    t2n += 0.5 * einsum("ijef,abef->ijab", tau, Wabef)
    x = einsum("imae,mbej->ijab", t2, Wmbej) - einsum("ie,ma,mbej->ijab", t1, t1, W[o, v, v, o])
    t2n += P(P(x, 0, 1), 2, 3)
    t2n += P(einsum("ie,abej->ijab", t1, W[v, v, v, o]), 0, 1)
    t2n -= P(einsum("ma,mbij->ijab", t1, W[o, v, o, o]), 2, 3)
    return t1n, t2n


def residual(C, G, Fmo, nocc, t1s, t2s):
    """Spin-orbital residuals (numerator - D*t) at given spin amplitudes."""
    W, f, no, ns = spin_integrals(C, G, Fmo, nocc)
    o, v = slice(0, no), slice(no, ns)
    fd = f - np.diag(np.diag(f))
    fo = np.diag(f)[o]
    fv = np.diag(f)[v]
    Dia = fo[:, None] - fv[None, :]
    Dijab = fo[:, None, None, None] + fo[None, :, None, None] - fv[None, None, :, None] - fv[None, None, None, :]
    n1, n2 = update(W, f, fd, t1s, t2s, o, v, no, ns - no)
    return n1 - Dia * t1s, n2 - Dijab * t2s


def spatial_to_spin(t1, t2):
    """Closed-shell spatial amplitudes -> spin-orbital amplitudes.

    t_{IA} = t_ia for equal spins; t_{IJAB} = t_ij^ab delta(sI,sA)delta(sJ,sB)
    - t_ij^ba delta(sI,sB)delta(sJ,sA)."""
    o, v = t1.shape
    no, nv = 2 * o, 2 * v
    si = np.arange(no) % 2
    pi = np.arange(no) // 2
    sa = np.arange(nv) % 2
    pa = np.arange(nv) // 2
    T1 = t1[pi[:, None], pa[None, :]] * (si[:, None] == sa[None, :])
    A = t2[pi[:, None, None, None], pi[None, :, None, None], pa[None, None, :, None], pa[None, None, None, :]]
    B = t2[pi[:, None, None, None], pi[None, :, None, None], pa[None, None, None, :], pa[None, None, :, None]]
    d1 = (si[:, None, None, None] == sa[None, None, :, None]) & (si[None, :, None, None] == sa[None, None, None, :])
    d2 = (si[:, None, None, None] == sa[None, None, None, :]) & (si[None, :, None, None] == sa[None, None, :, None])
    T2 = A * d1 - B * d2
    return T1, T2
