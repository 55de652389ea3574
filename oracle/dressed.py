"""T1-dressed, ZDO-factorised closed-shell CCSD residual as numpy GEMMs
(test infrastructure only).

This is the algorithm the HIP kernels implement, written with numpy so that
every kernel-level step can be compared on the CPU.  It is NOT a restatement
of the reference: the reference computes the contravariant residual of
proj-ccsd-spat-hf-1/2.f90 directly; this module computes the biorthogonal
residual Omega of the T1-transformed Hamiltonian (Koch, Christiansen,
Kobayashi, Jorgensen, Helgaker; Helgaker, Jorgensen, Olsen, "Molecular
Electronic-Structure Theory", sec. 13.7) and the tests pin the identities

    d1(i,a)   = Omega1[i,a]
    d2(i,j,a,b) = 2*Omega2[i,j,a,b] - Omega2[i,j,b,a]

against ``oracle.ccsd`` at random amplitudes.

Dressing with t1 (o x v):
    bra orbitals  Xo = Co,              Xv = Cv - Co t1
    ket orbitals  Yo = Co + Cv t1^T,    Yv = Cv
    g~(pq|rs) = sum_mn X_mp Y_mq gamma_mn X_nr Y_ns
"""
import numpy as np

einsum = lambda *a: np.einsum(*a, optimize=True)   # noqa: E731


def dress(C, o, t1):
    Co, Cv = C[:, :o], C[:, o:]
    X = {"o": Co, "v": Cv - Co @ t1}
    Y = {"o": Co + Cv @ t1.T, "v": Cv}
    return X, Y


def gt(X, Y, G, spec):
    """Dressed integral block g~_{pqrs} for occupation spec 'pqrs'."""
    A = einsum("mp,mq->mpq", X[spec[0]], Y[spec[1]])
    B = einsum("mr,ms->mrs", X[spec[2]], Y[spec[3]])
    return einsum("mpq,mn,nrs->pqrs", A, G, B)


def dressed_fock(C, F, G, o, t1):
    """F~ = X^T (Fao + 2 diag(G diag(dD)) - G*dD) Y with dD = Cv t1^T Co^T."""
    Co, Cv = C[:, :o], C[:, o:]
    X, Y = dress(C, o, t1)
    Fao = C @ F @ C.T
    dD = Cv @ t1.T @ Co.T
    M = Fao - G * dD
    M[np.diag_indices_from(M)] += 2.0 * (G @ np.diag(dD))
    Xf = np.concatenate([X["o"], X["v"]], axis=1)
    Yf = np.concatenate([Y["o"], Y["v"]], axis=1)
    return Xf.T @ M @ Yf


def omega(C, F, G, o, t1, t2):
    """Biorthogonal residual (Omega1[i,a], Omega2[i,j,a,b] = Omega_aibj)."""
    X, Y = dress(C, o, t1)
    Ft = dressed_fock(C, F, G, o, t1)
    Fov, Fvo = Ft[:o, o:], Ft[o:, :o]
    Foo, Fvv = Ft[:o, :o], Ft[o:, o:]
    g = lambda s: gt(X, Y, G, s)                       # noqa: E731
    u = 2 * t2 - t2.transpose(0, 1, 3, 2)              # u_ij^ab
    # singles
    w1 = einsum("kicd,adkc->ia", u, g("vvov"))         # A1: u_ki^cd g~_adkc
    w1 -= einsum("klac,kilc->ia", u, g("ooov"))        # B1: u_kl^ac g~_kilc
    w1 += einsum("ikac,kc->ia", u, Fov)                # C1: u_ik^ac F~_kc
    w1 += Fvo.T                                        # D1: F~_ai
    # doubles: A2 + B2
    gvovo = g("vovo")
    govov = g("ovov")                                  # undressed
    w2 = einsum("aibj->ijab", gvovo).copy()
    w2 += einsum("ijcd,acbd->ijab", t2, g("vvvv"))
    W = einsum("kilj->klij", g("oooo")) + einsum("ijcd,kcld->klij", t2, govov)
    w2 += einsum("klab,klij->ijab", t2, W)
    # C2, D2, E2 (then P_ij^ab)
    goovv = g("oovv")                                  # g~_kiac
    WC = einsum("kiac->aikc", goovv) - 0.5 * einsum("liad,kdlc->aikc", t2, govov)
    c2 = -0.5 * einsum("kjbc,aikc->ijab", t2, WC) - einsum("kibc,ajkc->ijab", t2, WC)
    L_aikc = 2 * g("voov") - einsum("acki->aikc", g("vvoo"))
    L_ldkc = 2 * govov - govov.transpose(0, 3, 2, 1)
    WD = L_aikc + 0.5 * einsum("ilad,ldkc->aikc", u, L_ldkc)
    d2 = 0.5 * einsum("jkbc,aikc->ijab", u, WD)
    Fvv_ = Fvv - einsum("klbd,ldkc->bc", u, govov)
    Foo_ = Foo + einsum("ljcd,kdlc->kj", u, govov)
    e2 = einsum("ijac,bc->ijab", t2, Fvv_) - einsum("ikab,kj->ijab", t2, Foo_)
    x = c2 + d2 + e2
    w2 += x + x.transpose(1, 0, 3, 2)
    return w1, w2


def contravariant(w2):
    """d2 = 2 Omega - Omega^T (swap of a and b)."""
    return 2 * w2 - w2.transpose(0, 1, 3, 2)
