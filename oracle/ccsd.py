"""Spatial-orbital closed-shell CCSD of the reference, restated term by term
(test infrastructure only).

Reference files (paths relative to the helios source tree):
  integrals   src/methods/coupledCluster/mod_cc.f90:2258-2277 (spat_hf_int)
  singles     src/methods/coupledCluster/ccsd/proj-ccsd-spat-hf-1.f90:14-119
  doubles     src/methods/coupledCluster/ccsd/proj-ccsd-spat-hf-2.f90:14-256
  guess       mod_cc.f90:634-650
  step        mod_cc.f90:987-1017
  DIIS        mod_cc.f90:1390-1423 (push), 1514-1555 + 1643-1671 (extrapolate),
              2133-2193 (dot products)
  energy      mod_cc.f90:1916-1952
  iteration   mod_cc.f90:820-891 and src/calculation/gIterator.f90:88-210

Index convention: numpy 0-based, occupied 0..o-1, virtual o..N-1; t1[i,a] and
t2[i,j,a,b] carry virtual indices counted from 0.  Each einsum below carries
the Fortran expression it restates.  The only liberty taken is the order of
floating-point summation (einsum instead of the reference's loop order).

Pinning (tests/test_oracle_ccsd.py):
  * two-electron molecules: CCSD equals the closed-form two-site FCI energy
    and the determinant FCI of ``oracle.fci``;
  * the converged energy equals that of the independent spin-orbital CCSD
    (``oracle.spin_ccsd``) to 1e-10 eV on several molecules;
  * residual identities at random amplitudes: the doubles residual equals
    2*Omega - Omega^T of the spin-orbital residual projected on closed-shell
    amplitudes (a term-level check that fails for any single dropped or
    mis-indexed term);
  * U -> 0 limit, size-extensivity for non-interacting fragments, invariance
    under occupied-occupied and virtual-virtual rotations.
"""
import numpy as np

einsum = lambda *a: np.einsum(*a, optimize=True)   # noqa: E731


# ----------------------------------------------------------------------------
# integrals
# ----------------------------------------------------------------------------
def eri(C, G):
    """(pq|rs) = sum_mn C_mp C_mq gamma_mn C_nr C_ns (ZDO)."""
    X = einsum("mp,mq->mpq", C, C)
    Y = einsum("mn,npq->mpq", G, X)
    return einsum("mpq,mrs->pqrs", X, Y)


def rint(C, G):
    """R(p,q,r,s) = 2(pq|rs) - (ps|rq)  (mod_cc.f90:2258-2277)."""
    g = eri(C, G)
    return 2.0 * g - g.transpose(0, 3, 2, 1)


class Blocks:
    """Sub-blocks of R by occupation string, e.g. B('voov') = R[v,o,o,v]."""

    def __init__(self, R, nocc):
        self.R = R
        self.o = nocc
        self.cache = {}

    def __call__(self, spec):
        if spec not in self.cache:
            sl = tuple(slice(0, self.o) if c == "o" else slice(self.o, None) for c in spec)
            self.cache[spec] = np.ascontiguousarray(self.R[sl])
        return self.cache[spec]


# ----------------------------------------------------------------------------
# projections
# ----------------------------------------------------------------------------
def singles(F, B, t1, t2):
    """d1(i,a) of proj-ccsd-spat-hf-1.f90:14-119."""
    o = t1.shape[0]
    Foo, Fov, Fvv = F[:o, :o], F[:o, o:], F[o:, o:]
    t2x = t2.transpose(1, 0, 2, 3)                       # t2(j,i,a,b) at [i,j,a,b]
    tau = t2 + einsum("ia,jb->ijab", t1, t1)
    d1 = Fov.copy()                                      # :20  F(i,a)
    d1 += einsum("aijb,jb->ia", B("voov"), t1)           # :27  R(a,i,j,b)t1(j,b)
    d1 += einsum("ab,ib->ia", Fvv, t1)                   # :37  F(a,b)t1(i,b)
    d1 -= einsum("ij,ja->ia", Foo, t1)                   # :46  F(i,j)t1(j,a)
    d1 += einsum("jb,ijab->ia", Fov, 2 * t2 - t2x)       # :56  F(j,b)(2t2(i,j,a,b)-t2(j,i,a,b))
    d1 += einsum("abjc,ijbc->ia", B("vvov"), tau)        # :68  R(a,b,j,c)(t2(i,j,b,c)+t1t1)
    d1 -= einsum("kbji,jkab->ia", B("ovoo"), tau)        # :81  R(k,b,j,i)(t2(j,k,a,b)+t1t1)
    d1 -= einsum("jb,ib,ja->ia", Fov, t1, t1)            # :93
    Rk = B("ovov")                                       # :106-110 R(k,c,j,b)(...)
    d1 += einsum("kcjb,kc,ijab->ia", Rk, t1, 2 * t2 - t2x)
    d1 -= einsum("kcjb,ib,jkac->ia", Rk, t1, t2)
    d1 -= einsum("kcjb,ka,jibc->ia", Rk, t1, t2)
    d1 -= einsum("kcjb,ka,jb,ic->ia", Rk, t1, t1, t1)
    return d1


def doubles_raw(F, B, t1, t2):
    """rez(i,j,a,b) of proj-ccsd-spat-hf-2.f90:23-251 for every (i,j,a,b)."""
    o = t1.shape[0]
    Foo, Fov, Fvv = F[:o, :o], F[:o, o:], F[o:, o:]
    tau = t2 + einsum("ia,jb->ijab", t1, t1)
    r = einsum("aibj->ijab", B("vovo")).copy()                              # :25
    Rvovv = B("vovv")                                                       # :31-32
    r += einsum("bjac,ic->ijab", Rvovv, t1) + einsum("aibc,jc->ijab", Rvovv, t1)
    Roovo = B("oovo")                                                       # :41-42
    r -= einsum("kibj,ka->ijab", Roovo, t1) + einsum("kjai,kb->ijab", Roovo, t1)
    r += einsum("kilj,klab->ijab", B("oooo"), tau)                          # :52 hole-hole ladder
    r += einsum("bdac,ijcd->ijab", B("vvvv"), tau)                          # :63 particle ladder
    # ring, :75-98
    Rvoov, Rovvo, Roovv, Rvvoo = B("voov"), B("ovvo"), B("oovv"), B("vvoo")
    r += 2 * einsum("ikac,bjkc->ijab", t2, Rvoov)
    r += 2 * einsum("jkbc,aikc->ijab", t2, Rvoov)
    r -= einsum("ikca,kcbj->ijab", t2, Rovvo)
    r -= einsum("jkcb,kcai->ijab", t2, Rovvo)
    r -= einsum("jkac,kcbi->ijab", t2, Rovvo)
    r -= einsum("ikbc,kcaj->ijab", t2, Rovvo)
    r -= einsum("kjac,kibc->ijab", t2, Roovv)
    r -= einsum("kibc,kjac->ijab", t2, Roovv)
    r -= einsum("ackj,ic,kb->ijab", Rvvoo, t1, t1)
    r -= einsum("bcki,jc,ka->ijab", Rvvoo, t1, t1)
    r -= einsum("kcbj,ic,ka->ijab", Rovvo, t1, t1)
    r -= einsum("kcai,jc,kb->ijab", Rovvo, t1, t1)
    # :112-126, sum over k,c,d
    Rovvv, Rvvov = B("ovvv"), B("vvov")
    r += einsum("kcbd,jd,ikac->ijab", Rovvv, t1, 2 * t2) - einsum("kcbd,jd,kiac->ijab", Rovvv, t1, t2)
    r -= einsum("kcbd,jkac,id->ijab", Rovvv, t2, t1)
    r -= einsum("kcbd,kjad,ic->ijab", Rovvv, t2, t1)
    r += einsum("kcad,id,jkbc->ijab", Rovvv, t1, 2 * t2) - einsum("kcad,id,kjbc->ijab", Rovvv, t1, t2)
    r -= einsum("kcad,ikbc,jd->ijab", Rovvv, t2, t1)
    r -= einsum("kcad,kibd,jc->ijab", Rovvv, t2, t1)
    t2s = 2 * t2 - t2.transpose(0, 1, 3, 2)              # 2t2(i,j,a,c)-t2(i,j,c,a) at [i,j,a,c]
    r += einsum("kd,kdbc,ijac->ijab", t1, Rovvv, t2s)
    r += einsum("kd,kdac,jibc->ijab", t1, Rovvv, t2s)
    r -= einsum("ackd,kb,ijcd->ijab", Rvvov, t1, tau)
    r -= einsum("bckd,ka,jicd->ijab", Rvvov, t1, tau)
    # :141-155, sum over k,c,l
    Rovoo, Rooov = B("ovoo"), B("ooov")
    r += einsum("kclj,lb,ikac->ijab", Rovoo, t1, -2 * t2) + einsum("kclj,lb,kiac->ijab", Rovoo, t1, t2)
    r += einsum("kclj,ikbc,la->ijab", Rovoo, t2, t1)
    r += einsum("kclj,ilcb,ka->ijab", Rovoo, t2, t1)
    r += einsum("kcli,la,jkbc->ijab", Rovoo, t1, -2 * t2) + einsum("kcli,la,kjbc->ijab", Rovoo, t1, t2)
    r += einsum("kcli,jkac,lb->ijab", Rovoo, t2, t1)
    r += einsum("kcli,jlca,kb->ijab", Rovoo, t2, t1)
    r += einsum("lc,kjlc,ikab->ijab", t1, Rooov, -2 * t2) + einsum("lc,kjlc,kiab->ijab", t1, Rooov, t2)
    r += einsum("lc,kilc,jkba->ijab", t1, Rooov, -2 * t2) + einsum("lc,kilc,kjba->ijab", t1, Rooov, t2)
    r += einsum("kclj,ic,klab->ijab", Rovoo, t1, tau)
    r += einsum("kcli,jc,klba->ijab", Rovoo, t1, tau)
    # :171-199, sum over k,c,l,d of R(k,c,l,d)*(...)
    Rk = B("ovov")

    def e(s, *ops):
        return einsum("kcld," + s + "->ijab", Rk, *ops)

    m = (e("ljdb,ic,ka", t2, t1, t1) + e("lida,jc,kb", t2, t1, t1)          # :173
         + e("ljab,kc,id", t2, t1, t1) + e("liba,kc,jd", t2, t1, t1)        # :174
         + e("ijac,kb,ld", t2, t1, t1) + e("jibc,ka,ld", t2, t1, t1)        # :175
         + e("kicd,ljab", t2, t2) + e("kjcd,liba", t2, t2)                  # :176
         + e("ijac,klbd", t2, t2) + e("jibc,klad", t2, t2)                  # :177
         + e("ikca,ljdb", t2, t2) + e("jkcb,lida", t2, t2)                  # :178
         - e("ikac,ljdb", t2, t2) - e("jkbc,lida", t2, t2))                 # :179
    r -= 2 * m
    r += e("jldb,ic,ka", t2, t1, t1) + e("ilda,jc,kb", t2, t1, t1)          # :182
    r += e("lidb,jc,ka", t2, t1, t1) + e("ljda,ic,kb", t2, t1, t1)          # :183
    r += e("ilcb,jd,ka", t2, t1, t1) + e("jlca,id,kb", t2, t1, t1)          # :184
    r += e("liab,kc,jd", t2, t1, t1) + e("ljba,kc,id", t2, t1, t1)          # :185
    r += e("ijca,kb,ld", t2, t1, t1) + e("jicb,ka,ld", t2, t1, t1)          # :186
    r += e("klab,ic,jd", t2, t1, t1)                                        # :188
    r += e("ijcd,klab", t2, tau)                                            # :189
    r += e("ic,ka,jd,lb", t1, t1, t1, t1)                                   # :190
    r += e("kjcd,liab", t2, t2) + e("kicd,ljba", t2, t2)                    # :192
    r += e("ijbc,klad", t2, t2) + e("jiac,klbd", t2, t2)                    # :193
    r -= e("lidb,jkac", t2, t2) - e("lidb,jkca", t2, t2)                    # :194
    r -= e("ljda,ikbc", t2, t2) - e("ljda,ikcb", t2, t2)                    # :195
    r += 0.5 * (e("ikca,jldb", t2, t2) + e("jkcb,ilda", t2, t2)             # :196-197
                + e("jkda,ilcb", t2, t2) + e("ikdb,jlca", t2, t2))
    # Fock, :212-213 and :222-223
    r += einsum("ca,ijcb->ijab", Fvv, 2 * t2) - einsum("ca,jicb->ijab", Fvv, t2)
    r += einsum("cb,jica->ijab", Fvv, 2 * t2) - einsum("cb,ijca->ijab", Fvv, t2)
    r -= einsum("ki,kjab->ijab", Foo, 2 * t2) - einsum("ki,kjba->ijab", Foo, t2)
    r -= einsum("kj,kiba->ijab", Foo, 2 * t2) - einsum("kj,kiab->ijab", Foo, t2)
    # F(k,c) t1 t2, :233-246
    r -= 2 * (einsum("kc,kjab,ic->ijab", Fov, t2, t1) + einsum("kc,kiba,jc->ijab", Fov, t2, t1)
              + einsum("kc,ijac,kb->ijab", Fov, t2, t1) + einsum("kc,jibc,ka->ijab", Fov, t2, t1))
    r += (einsum("kc,kiab,jc->ijab", Fov, t2, t1) + einsum("kc,kjba,ic->ijab", Fov, t2, t1)
          + einsum("kc,jiac,kb->ijab", Fov, t2, t1) + einsum("kc,ijbc,ka->ijab", Fov, t2, t1))
    return r


def pair_index(o, v):
    """mm(i,a) = i*v + a: excSet order, i-major a-minor (mod_cc.f90:419-424)."""
    return np.arange(o)[:, None] * v + np.arange(v)[None, :]


def mirror(r):
    """d2(i,j,a,b)=rez for mm(i,a)<=mm(j,b), d2(j,i,b,a)=rez (proj-ccsd-spat-hf-2.f90:253)."""
    o, _, v, _ = r.shape
    mm = pair_index(o, v)
    upper = mm[:, None, :, None] <= mm[None, :, None, :]      # [i,j,a,b]: mm(i,a)<=mm(j,b)
    return np.where(upper, r, r.transpose(1, 0, 3, 2))


def doubles(F, B, t1, t2):
    return mirror(doubles_raw(F, B, t1, t2))


# ----------------------------------------------------------------------------
# guess / step / energy
# ----------------------------------------------------------------------------
def denominators(F, o):
    e = np.diag(F)
    eo, ev = e[:o], e[o:]
    D1 = eo[:, None] - ev[None, :]
    D2 = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    return D1, D2


def guess(F, B, o):
    """t1=F_ia/(F_ii-F_aa); t2(i,j,a,b)=R(i,a,j,b)/(F_ii+F_jj-F_aa-F_bb)
    (mod_cc.f90:634-650; the contravariant MP2 amplitude)."""
    D1, D2 = denominators(F, o)
    t1 = F[:o, o:] / D1
    t2 = mirror(einsum("iajb->ijab", B("ovov")) / D2)
    return t1, t2


def mp2_amplitudes(C, G, E, o):
    """True MP2 amplitudes (ia|jb)/(e_i+e_j-e_a-e_b) (mod_mbpt.f90:136-165)."""
    g = eri(C, G)
    eo, ev = E[:o], E[o:]
    D2 = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    return einsum("iajb->ijab", g[:o, o:, :o, o:]) / D2


def mp2_energy(C, G, E, o):
    """E2 = sum (2(ia|jb)-(ib|ja))(ia|jb)/D  (mod_mbpt.f90:131-175)."""
    g = eri(C, G)
    V = einsum("iajb->ijab", g[:o, o:, :o, o:])
    eo, ev = E[:o], E[o:]
    D2 = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    return float(((2 * V - V.transpose(0, 1, 3, 2)) * V / D2).sum())


def step(F, t1, t2, d1, d2, s1=1.0, s2=1.0, projtype=2):
    """stepCC for r-ccsd (mod_cc.f90:987-1017).  Returns new t1, t2 and
    accuracy(1:2) = max|d1|, max|d2|."""
    o = t1.shape[0]
    D1, D2 = denominators(F, o)
    t1n = t1 + s1 * d1 / D1
    if projtype == 2:
        num = 2 * d2 + d2.transpose(0, 1, 3, 2)
    else:
        num = d2
    t2n = mirror(t2 + s2 * num / D2)
    return t1n, t2n, (float(np.abs(d1).max()), float(np.abs(d2).max()))


def energy(F, B, t1, t2, eref):
    """energyCC r-ccsd: energy(1..5) (mod_cc.f90:1916-1952)."""
    o, v = t1.shape
    sum1 = float((F[:o, o:] * t1).sum())
    Rv = einsum("aibj->ijab", B("vovo"))
    tau = t2 + einsum("ia,jb->ijab", t1, t1)
    prod = Rv * tau
    mm = pair_index(o, v)
    lt = mm[:, None, :, None] < mm[None, :, None, :]
    eq = mm[:, None, :, None] == mm[None, :, None, :]
    sum2 = float(prod[lt].sum())
    sum3 = float(prod[eq].sum())
    e = np.zeros(5)
    e[0] = 2 * sum1 + (2 * sum2 + sum3) + eref
    e[1] = eref
    e[2] = 2 * sum1
    e[3] = 2 * sum2 + sum3
    return e


# ----------------------------------------------------------------------------
# DIIS (mod_cc.f90:1390-1423, 1514-1555, 1643-1671, 2133-2193)
# ----------------------------------------------------------------------------
class DIIS:
    """History and extrapolation of the reference.  ``scale=True`` divides B by
    its largest diagonal element before the bordered eigen-solve; the
    coefficients are unchanged in exact arithmetic, but the solve stays
    accurate when B's entries fall far below the border's 1's (residuals under
    ~1e-8), where the unscaled form loses all digits."""

    def __init__(self, o, v, nd, scale=False):
        self.o, self.v, self.nd, self.scale = o, v, nd, scale
        ne = o * v
        self.np2 = ne * (ne + 1) // 2
        self.st1 = np.zeros((ne, nd))
        self.sd1 = np.zeros((ne, nd))
        self.st2 = np.zeros((self.np2, nd))
        self.sd2 = np.zeros((self.np2, nd))
        self.B = np.zeros((nd + 1, nd + 1))
        mm = pair_index(o, v).ravel()
        # packed order: mm = 1..Ne, nn = mm..Ne
        ii, aa = np.divmod(np.arange(ne), v)
        rows, cols = np.triu_indices(ne)
        self.pi, self.pa = ii[rows], aa[rows]
        self.pj, self.pb = ii[cols], aa[cols]
        assert (mm == np.arange(ne)).all()

    def pack(self, x):
        return x[self.pi, self.pj, self.pa, self.pb]

    def unpack(self, p):
        o, v = self.o, self.v
        t = np.zeros((o, o, v, v))
        t[self.pi, self.pj, self.pa, self.pb] = p
        t[self.pj, self.pi, self.pb, self.pa] = p
        return t

    def push(self, t1, t2, d1, d2, projtype=2):
        self.st1[:, 1:] = self.st1[:, :-1].copy()
        self.sd1[:, 1:] = self.sd1[:, :-1].copy()
        self.st2[:, 1:] = self.st2[:, :-1].copy()
        self.sd2[:, 1:] = self.sd2[:, :-1].copy()
        self.sd1[:, 0] = d1.ravel()
        self.st1[:, 0] = t1.ravel()
        self.st2[:, 0] = self.pack(t2)
        self.sd2[:, 0] = self.pack(2 * d2 + d2.transpose(0, 1, 3, 2) if projtype == 2 else d2)

    def extrapolate(self, it):
        """newCCVectors(iteration=it) -> (t1, t2) or None when it == 1."""
        B = self.B
        B[1:it, 1:it] = B[0:it - 1, 0:it - 1].copy()
        for k in range(it):
            B[k, 0] = B[0, k] = self.sd1[:, k] @ self.sd1[:, 0] + self.sd2[:, k] @ self.sd2[:, 0]
        if it == 1:
            return None
        ub = it + 1
        B[:ub, it] = 1.0
        B[it, :ub] = 1.0
        B[it, it] = 0.0
        M = B[:ub, :ub].copy()
        if self.scale:
            M[:it, :it] /= np.max(np.diag(M)[:it])
        lam, V = np.linalg.eigh(M)
        c = (V[:it, :] * V[it, :][None, :] / lam[None, :]).sum(1)
        self.coef = c          # newest vector first, as stored
        t1 = (self.st1[:, :it] @ c).reshape(self.o, self.v)
        t2 = self.unpack(self.st2[:, :it] @ c)
        return t1, t2


# ----------------------------------------------------------------------------
# solver
# ----------------------------------------------------------------------------
class CCSD:
    """State of one r-ccsd solve: initCC + iterationCC + energyCC."""

    def __init__(self, C, Fmo, G, nocc, eref, diis_steps=20, diis=True,
                 s1=1.0, s2=1.0, projtype=2, diis_scale=False):
        self.C, self.F, self.G, self.o, self.eref = C, Fmo, G, nocc, eref
        self.v = C.shape[1] - nocc
        self.B = Blocks(rint(C, G), nocc)
        self.t1, self.t2 = guess(Fmo, self.B, nocc)
        self.diis = DIIS(nocc, self.v, diis_steps, diis_scale) if diis else None
        self.nd = diis_steps
        self.s1, self.s2, self.projtype = s1, s2, projtype

    def residuals(self, t1=None, t2=None):
        t1 = self.t1 if t1 is None else t1
        t2 = self.t2 if t2 is None else t2
        return singles(self.F, self.B, t1, t2), doubles(self.F, self.B, t1, t2)

    def iterate(self, iteration, eps):
        """iterationCC (mod_cc.f90:820-891); returns accuracy(5)."""
        d1, d2 = self.residuals()
        self.t1, self.t2, acc = step(self.F, self.t1, self.t2, d1, d2,
                                     self.s1, self.s2, self.projtype)
        accuracy = np.array([acc[0], acc[1], -1.0, -1.0, -1.0])
        if max(accuracy) < eps:
            return accuracy
        if self.diis is not None:
            self.diis.push(self.t1, self.t2, d1, d2, self.projtype)
            res = self.diis.extrapolate(min(iteration, self.nd))
            if res is not None:
                self.t1, self.t2 = res
        return accuracy

    def energy(self):
        return energy(self.F, self.B, self.t1, self.t2, self.eref)


def iterator(callsub, energysub, maxiter, eps, divergence=1e3, stagnation_after=200,
             stagnation_percent=60.0, check_divergence=True, check_stagnation=True,
             stop_flag=None):
    """gIterator.f90:88-210.  Returns (status, iterations, energy(5), history).

    status: 1 converged, -1 max iterations, -2 diverged, -3 stagnated,
    -4 interrupted (mod_helios.f90:119-120)."""
    it = 0
    successful = 1
    cur, prev = 0.0, 10.0
    hist = []
    while True:
        it += 1
        acc = callsub(it, eps)
        en = energysub()
        prev, cur = cur, float(np.max(acc))
        if cur < prev:
            successful += 1
        hist.append((cur, en[0]))
        if stop_flag is not None and stop_flag(it):
            return -4, it, en, hist
        if cur < eps:
            return 1, it, en, hist
        if check_divergence and cur > divergence:
            return -2, it, en, hist
        if it + 1 > maxiter:
            return -1, it, en, hist
        if check_stagnation and it > stagnation_after and 100.0 * successful / it <= stagnation_percent:
            return -3, it, en, hist


def solve(C, Fmo, G, nocc, eref, eps=1e-10, maxiter=2000, **kw):
    cc = CCSD(C, Fmo, G, nocc, eref, **kw)
    status, it, en, hist = iterator(cc.iterate, cc.energy, maxiter, eps)
    return dict(status=status, iterations=it, energy=en, t1=cc.t1, t2=cc.t2, history=hist, cc=cc)
