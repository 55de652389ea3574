"""The (T) energy terms of the reference's spin-orbital CCSD(T), restated
(test infrastructure only): energyCC('spin-r-ccsd(t)'),
src/methods/coupledCluster/mod_cc.f90:2028-2111, fed with the closed-shell
amplitudes converted to spin orbitals as convertSpatialToSpinCC does
(mod_cc.f90:2601-2678), the antisymmetrised spin-orbital integrals of
spin_hf_int (mod_cc.f90:2283-2308: R(a,b,c,d) = [ab|cd] - [ad|bc] with spin
conservation by index parity) and the spin-orbital Fock matrix of
prepareFock('spin') (F(2i-1,2j-1) = F(2i,2j) = F(i,j)).

Spin orbitals are numbered as the reference's: spatial p gives 2p (alpha)
and 2p+1 (beta) here (0-based), occupied spin orbitals first.  The loop is
the reference's six-index one, vectorised over (i,j,k) for each (a,b,c):

  sum1 = sum_l [ R(j,c,k,l) t(i,l,a,b) - R(i,c,k,l) t(j,l,a,b) - R(j,c,i,l) t(k,l,a,b)
               - R(j,a,k,l) t(i,l,c,b) + R(i,a,k,l) t(j,l,c,b) + R(j,a,i,l) t(k,l,c,b)
               - R(j,b,k,l) t(i,l,a,c) + R(i,b,k,l) t(j,l,a,c) + R(j,b,i,l) t(k,l,a,c) ]
  sum2 = sum_d [ R(b,k,c,d) t(i,j,a,d) - R(b,i,c,d) t(k,j,a,d) - R(b,j,c,d) t(i,k,a,d)
               - R(a,k,c,d) t(i,j,b,d) + R(a,i,c,d) t(k,j,b,d) + R(a,j,c,d) t(i,k,b,d)
               - R(b,k,a,d) t(i,j,c,d) + R(b,i,a,d) t(k,j,c,d) + R(b,j,a,d) t(i,k,c,d) ]
  amp  = (sum1 - sum2) / denom,  denom = F(a,a)+F(b,b)+F(c,c)-F(i,i)-F(j,j)-F(k,k)
  c1 = sum amp^2 denom / 36,  c2 = sum amp t1(i,a) R(b,j,c,k) / 4,  c3 = sum amp t2(i,j,a,b) F(k,c)

over index sets whose occupied and virtual parities agree.  energy(1) =
eref + one + two + c1 + c2 + c3 and energy(5) = c1.
"""
import numpy as np

einsum = lambda *a: np.einsum(*a, optimize=True)   # noqa: E731


def spin_integrals(C, G):
    """R(p,q,r,s) = [pq|rs] - [ps|qr] over 2n spin orbitals (spin_hf_int)."""
    n = C.shape[0]
    X = einsum("mp,mq->mpq", C, C)
    g = einsum("mpq,mn,nrs->pqrs", X, G, X)          # spatial (pq|rs)
    s = np.arange(2 * n) % 2                          # spin of each spin orbital
    sp = np.arange(2 * n) // 2                        # its spatial orbital
    gs = g[np.ix_(sp, sp, sp, sp)]
    same = lambda a, b: (s[a][:, None] == s[b][None, :])   # noqa: E731
    idx = np.arange(2 * n)
    m1 = same(idx, idx)[:, :, None, None] & same(idx, idx)[None, None, :, :]      # [ab|cd]: a~b, c~d
    m2 = same(idx, idx)[:, None, None, :] & same(idx, idx)[None, :, :, None]      # [ad|bc]: a~d, b~c
    return np.where(m1, gs, 0.0) - np.where(m2, gs.transpose(0, 2, 3, 1), 0.0)   # (ad|bc) at [a,b,c,d]


def spin_amplitudes(t1, t2):
    """convertSpatialToSpinCC: st1 (2o x 2v), st2 (2o,2o,2v,2v)."""
    o, v = t1.shape
    st1 = np.zeros((2 * o, 2 * v))
    st1[0::2, 0::2] = t1
    st1[1::2, 1::2] = t1
    st2 = np.zeros((2 * o, 2 * o, 2 * v, 2 * v))
    ab = t2                                            # alpha(i) beta(j) -> alpha(a) beta(b)
    st2[0::2, 1::2, 0::2, 1::2] = ab
    st2[1::2, 0::2, 0::2, 1::2] = -ab.transpose(1, 0, 2, 3)
    st2[0::2, 1::2, 1::2, 0::2] = -ab.transpose(0, 1, 3, 2)
    st2[1::2, 0::2, 1::2, 0::2] = ab.transpose(1, 0, 3, 2)
    aa = t2 - t2.transpose(1, 0, 2, 3)
    st2[1::2, 1::2, 1::2, 1::2] = aa
    st2[0::2, 0::2, 0::2, 0::2] = aa
    return st1, st2


def triples(C, F, G, nocc, t1, t2):
    """(c1, c2, c3) of the reference's (T) loop for the closed-shell amplitudes."""
    n = C.shape[0]
    no, ns = 2 * nocc, 2 * n
    R = spin_integrals(C, G)
    Fs = np.zeros((ns, ns))
    Fs[0::2, 0::2] = F
    Fs[1::2, 1::2] = F
    st1, st2 = spin_amplitudes(t1, t2)
    e = np.diag(Fs)
    o_, v_ = slice(0, no), slice(no, ns)
    Rooov = R[o_, v_, o_, o_]      # R(j,c,k,l): occ, vir, occ, occ
    Rvovv = R[v_, o_, v_, v_]      # R(b,k,c,d): vir, occ, vir, vir
    Rvovo = R[v_, o_, v_, o_]      # R(b,j,c,k)
    par_o = np.arange(no) % 2
    par_v = (np.arange(no, ns)) % 2
    pijk = (par_o[:, None, None] + par_o[None, :, None] + par_o[None, None, :]) % 2
    eo = e[:no]
    eijk = eo[:, None, None] + eo[None, :, None] + eo[None, None, :]
    nv = ns - no
    c1 = c2 = c3 = 0.0
    T = st2                         # t(i,j,a,b), virtual indices from 0
    Fov = Fs[:no, no:]
    for a in range(nv):
        for b in range(nv):
            for c in range(nv):
                pabc = (par_v[a] + par_v[b] + par_v[c]) % 2
                mask = pijk == pabc
                if not mask.any():
                    continue
                s1 = (einsum("jkl,il->ijk", Rooov[:, c, :, :], T[:, :, a, b])
                      - einsum("ikl,jl->ijk", Rooov[:, c, :, :], T[:, :, a, b])
                      - einsum("jil,kl->ijk", Rooov[:, c, :, :], T[:, :, a, b])
                      - einsum("jkl,il->ijk", Rooov[:, a, :, :], T[:, :, c, b])
                      + einsum("ikl,jl->ijk", Rooov[:, a, :, :], T[:, :, c, b])
                      + einsum("jil,kl->ijk", Rooov[:, a, :, :], T[:, :, c, b])
                      - einsum("jkl,il->ijk", Rooov[:, b, :, :], T[:, :, a, c])
                      + einsum("ikl,jl->ijk", Rooov[:, b, :, :], T[:, :, a, c])
                      + einsum("jil,kl->ijk", Rooov[:, b, :, :], T[:, :, a, c]))
                s2 = (einsum("kd,ijd->ijk", Rvovv[b, :, c, :], T[:, :, a, :])
                      - einsum("id,kjd->ijk", Rvovv[b, :, c, :], T[:, :, a, :])
                      - einsum("jd,ikd->ijk", Rvovv[b, :, c, :], T[:, :, a, :])
                      - einsum("kd,ijd->ijk", Rvovv[a, :, c, :], T[:, :, b, :])
                      + einsum("id,kjd->ijk", Rvovv[a, :, c, :], T[:, :, b, :])
                      + einsum("jd,ikd->ijk", Rvovv[a, :, c, :], T[:, :, b, :])
                      - einsum("kd,ijd->ijk", Rvovv[b, :, a, :], T[:, :, c, :])
                      + einsum("id,kjd->ijk", Rvovv[b, :, a, :], T[:, :, c, :])
                      + einsum("jd,ikd->ijk", Rvovv[b, :, a, :], T[:, :, c, :]))
                denom = e[no + a] + e[no + b] + e[no + c] - eijk
                amp = np.where(mask, (s1 - s2) / denom, 0.0)
                c1 += float((amp * amp * denom).sum())
                c2 += float(einsum("ijk,i,jk->", amp, st1[:, a], Rvovo[b, :, c, :]))
                c3 += float(einsum("ijk,ij,k->", amp, T[:, :, a, b], Fov[:, c]))
    return c1 / 36.0, c2 / 4.0, c3


def closed_shell(C, F, G, nocc, t1, t2):
    """E[T] and E[ST] of the standard (T) correction from the closed-shell
    amplitudes, written directly in spatial orbitals (no reference
    counterpart; pinned against ``triples`` above, whose c1, c2 are their
    negatives):

      w_ijk^abc = sum_d (bd|ai) t_kj^cd - sum_l (ck|jl) t_il^ab
      W_ijk^abc = sum over the 6 simultaneous permutations of (ia), (jb), (kc) of w
      V_ijk^abc = t_i^a (bj|ck) + t_j^b (ai|ck) + t_k^c (ai|bj)
      R3(X)_abc = 4 X_abc + X_bca + X_cab - 2 X_acb - 2 X_bac - 2 X_cba
      E[T] = sum W R3(W) / (3 D),  E[ST] = sum V R3(W) / (3 D),
      D = e_i + e_j + e_k - e_a - e_b - e_c.
    """
    o = nocc
    X = einsum("mp,mq->mpq", C, C)
    g = einsum("mpq,mn,nrs->pqrs", X, G, X)
    e = np.diag(F)
    eo, ev = e[:o], e[o:]
    w = einsum("bdai,kjcd->ijkabc", g[o:, o:, o:, :o], t2) - einsum("ckjl,ilab->ijkabc", g[o:, :o, :o, :o], t2)
    W = (w + w.transpose(0, 2, 1, 3, 5, 4) + w.transpose(1, 0, 2, 4, 3, 5) + w.transpose(1, 2, 0, 4, 5, 3)
         + w.transpose(2, 0, 1, 5, 3, 4) + w.transpose(2, 1, 0, 5, 4, 3))
    gv = g[o:, :o, o:, :o]
    V = (einsum("ia,bjck->ijkabc", t1, gv) + einsum("jb,aick->ijkabc", t1, gv)
         + einsum("kc,aibj->ijkabc", t1, gv))
    D = (eo[:, None, None, None, None, None] + eo[None, :, None, None, None, None]
         + eo[None, None, :, None, None, None] - ev[None, None, None, :, None, None]
         - ev[None, None, None, None, :, None] - ev[None, None, None, None, None, :])
    R3 = 4 * W + W.transpose(0, 1, 2, 4, 5, 3) + W.transpose(0, 1, 2, 5, 3, 4) - 2 * W.transpose(0, 1, 2, 3, 5, 4) \
        - 2 * W.transpose(0, 1, 2, 4, 3, 5) - 2 * W.transpose(0, 1, 2, 5, 4, 3)
    return float((W * R3 / (3 * D)).sum()), float((V * R3 / (3 * D)).sum())


def closed_shell_sliced(C, F, G, nocc, t1, t2):
    """``closed_shell`` one occupied index i at a time (o^2 v^3 memory instead
    of o^3 v^3), for the larger parity cases.  For fixed i, W_ijk^abc needs w
    with i in each of its three occupied positions:

      A0[j,k,a,b,c] = w_ijk^abc,  A1[x,z,a,b,c] = w_xiz^abc,  A2[x,y,a,b,c] = w_xyi^abc

    and the six simultaneous permutations read
      w_ijk^abc = A0[j,k,a,b,c]   w_ikj^acb = A0[k,j,a,c,b]   w_jik^bac = A1[j,k,b,a,c]
      w_jki^bca = A2[j,k,b,c,a]   w_kij^cab = A1[k,j,c,a,b]   w_kji^cba = A2[k,j,c,b,a].
    """
    o = nocc
    X = einsum("mp,mq->mpq", C, C)
    g = einsum("mpq,mn,nrs->pqrs", X, G, X)
    e = np.diag(F)
    eo, ev = e[:o], e[o:]
    gvvvo = np.ascontiguousarray(g[o:, o:, o:, :o])    # (bd|ax)
    gvooo = np.ascontiguousarray(g[o:, :o, :o, :o])    # (cz|yl)
    gv = np.ascontiguousarray(g[o:, :o, o:, :o])       # (ai|bj)
    Dv = -(ev[:, None, None] + ev[None, :, None] + ev[None, None, :])
    eT = eST = 0.0
    for i in range(o):
        A0 = einsum("bda,zycd->yzabc", gvvvo[..., i], t2) - einsum("czyl,lab->yzabc", gvooo, t2[i])
        A1 = einsum("bdax,zcd->xzabc", gvvvo, t2[:, i]) - einsum("czl,xlab->xzabc", gvooo[:, :, i, :], t2)
        A2 = einsum("bdax,ycd->xyabc", gvvvo, t2[i]) - einsum("cyl,xlab->xyabc", gvooo[:, i], t2)
        W = (A0 + A0.transpose(1, 0, 2, 4, 3) + A1.transpose(0, 1, 3, 2, 4) + A2.transpose(0, 1, 4, 2, 3)
             + A1.transpose(1, 0, 3, 4, 2) + A2.transpose(1, 0, 4, 3, 2))
        V = (einsum("a,bjck->jkabc", t1[i], gv) + einsum("jb,ack->jkabc", t1, gv[:, i])
             + einsum("kc,abj->jkabc", t1, gv[:, i]))
        R3 = 4 * W + W.transpose(0, 1, 3, 4, 2) + W.transpose(0, 1, 4, 2, 3) - 2 * W.transpose(0, 1, 2, 4, 3) \
            - 2 * W.transpose(0, 1, 3, 2, 4) - 2 * W.transpose(0, 1, 4, 3, 2)
        D = (e[i] + eo[:, None, None, None, None] + eo[None, :, None, None, None]) + Dv[None, None]
        eT += float((W * R3 / (3 * D)).sum())
        eST += float((V * R3 / (3 * D)).sum())
    return eT, eST
