"""Singlet LR-CCSD (EOM-CCSD) excitation energies of a converged batch.

The reference's LR-CCSD (src/methods/coupledCluster/lrccsd/mod_lrccsd.f90)
follows one state at a time: ``iterationLR`` (:176-212) applies the CCSD
Jacobian to the trial vector (projections proj-lrccsd-spin-hf-1/2.f90 on
spin orbitals), ``stepLR`` (:216-326) forms the Rayleigh quotient
omega = <d, r>/<r, r>, steps r -= step (d - omega r), keeps the singlet
combination for spin = 0 and normalises, and DIIS accelerates the sequence;
at convergence omega is an eigenvalue of the CCSD Jacobian -- the LR-CCSD
excitation energy -- and ``energyLR`` (:330-389) reports it.

Here the same eigenvalues come from the closed-shell (singlet) Jacobian
J = dOmega/dt at the converged amplitudes, with a block Davidson solve for
the ``nroots`` lowest roots of every molecule of the batch at once:
Rayleigh-Ritz on an orthonormal subspace, residual corrections
preconditioned by the orbital-energy differences, complex Ritz pairs kept in
the working set, restarts onto the current and previous Ritz vectors.  The
subspace lives in device memory (``hcc_lr_space``): the Jacobian products
(``hcc_lr_sigma_slots``: central differences of the FP64 residual), the
dot products, linear combinations and the preconditioner run as batched
kernels over all molecules, and the host solves only the small projected
eigenproblems and decides what to combine next.  Trial vectors start as the
lowest CIS singlets (the reference's default ``cis`` guess) or as unit
singles (its ``unt``).  The difference step h = 1e-5 keeps sigma within
about 1e-9 of J r (the truncation error is O(h^2) times the third
derivative of a quartic polynomial; the FP64 rounding error is about
1e-16 |Omega| / h).  Triplets (spin = 1) need the spin-orbital Jacobian and
are not provided.
"""
import numpy as np
import scipy.linalg
from threadpoolctl import threadpool_limits


def cis_singlets(C, gamma, e, o, count=None):
    """Singlet CIS of the RHF reference: eigenvalues (ascending) and
    eigenvectors (columns over (i, a), i slowest) of
    A[(i a), (j b)] = (e_a - e_i) d_ij d_ab + 2 [ai|jb] - [ab|ji],
    [pq|rs] = sum_mn C_mp C_mq gamma_mn C_nr C_ns -- the matrix the reference
    diagonalises for its 'cis' guess (mod_lrccsd.f90:1253-1282 with
    spat_hf_int, :1423-1445).  With ``count``, only the lowest ``count``
    pairs, from LAPACK's MRRR solver on one thread (two to three times faster
    than the full spectrum at o v of a few hundred; several BLAS threads only
    compete with each other at these sizes)."""
    n = C.shape[0]
    v = n - o
    Co, Cv = C[:, :o], C[:, o:]
    X = (Co[:, :, None] * Cv[:, None, :]).reshape(n, o * v)       # C_mi C_ma
    coul = X.T @ gamma @ X                                          # [ai|jb]
    Y = (Cv[:, :, None] * Cv[:, None, :]).reshape(n, v * v)       # C_ma C_mb
    Z = (Co[:, :, None] * Co[:, None, :]).reshape(n, o * o)       # C_mi C_mj
    exch = (Y.T @ gamma @ Z).reshape(v, v, o, o).transpose(2, 0, 3, 1).reshape(o * v, o * v)   # [ab|ij]
    A = 2.0 * coul - exch
    A[np.diag_indices(o * v)] += (e[o:][None, :] - e[:o][:, None]).ravel()
    if count is None or count >= o * v:
        return np.linalg.eigh(A)
    with threadpool_limits(1):
        return scipy.linalg.eigh(A, subset_by_index=[0, count - 1], driver="evr")


def average_excitation_level(r1, r2):
    """The reference's average excitation level of a state,
    (S1 + 2 S2) / (S1 + S2) with S1 = sum r1^2 and S2 = sum_{i<j, a<b} r2^2 over
    spin orbitals (mod_lrccsd.f90:340-363), for a singlet given by its
    closed-shell components r1[i, a], r2[i, j, a, b] (pair-symmetric).  The
    spin-orbital amplitudes follow from the closed-shell ones as the
    amplitudes do (convertSpatialToSpinCC): alpha and beta singles equal r1;
    the four mixed-spin doubles blocks each hold r2 up to index order and sign,
    the two same-spin blocks r2 - r2(j i a b); each (i<j, a<b) element occurs
    four times in the antisymmetric tensor."""
    s1 = 2.0 * float(np.sum(r1 * r1))
    aa = r2 - r2.transpose(1, 0, 2, 3)
    s2 = (4.0 * float(np.sum(r2 * r2)) + 2.0 * float(np.sum(aa * aa))) / 4.0
    return (s1 + 2.0 * s2) / (s1 + s2)


def ovov_integrals(C, gamma, o):
    """(ia|jb) at [i, a, j, b] (ZDO: sum_mn C_mi C_ma gamma_mn C_nj C_nb)."""
    n = C.shape[0]
    v = n - o
    X = (C[:, :o, None] * C[:, None, o:]).reshape(n, o * v)
    return (X.T @ gamma @ X).reshape(o, v, o, v)


def reference_weight(g, f_ov, t1, omega, r1, r2):
    """The reference's r0 of a state (energyLR, mod_lrccsd.f90:365-379):
    r0 = (sum_ia Fbar_ia r_ia + 1/4 sum_ijab <ij||ab> r_ijab) / omega over
    spin orbitals, with Fbar_ia = F_ia + sum_kd t_kd <ik||ad>
    (proj-lrccsd-spin-hf-intermediates.f90:369-382), for the vector normalised
    as the reference's ``normalize`` does (sum r1^2 + sum_{i<j, a<b} r2^2 = 1);
    its magnitude is returned (an eigenvector's sign is arbitrary: the
    reference's follows from its iteration history).  For a
    singlet in closed-shell components (r1, r2 pair-symmetric) the spin sums
    reduce to 2 sum_ia Fbar_ia r1_ia with Fbar_ia = F_ia + sum_kd t1_kd
    (2 (ia|kd) - (id|ka)), and sum_ijab (2 (ia|jb) - (ib|ja)) r2_ijab (the
    closed-shell form of 1/4 sum <ij||ab> r_ijab, as in the CCSD energy).
    g: (ia|jb) at [i, a, j, b]; f_ov: the MO Fock block F[i, o + a]."""
    aa = r2 - r2.transpose(1, 0, 2, 3)
    s1 = 2.0 * float(np.sum(r1 * r1))
    s2 = (4.0 * float(np.sum(r2 * r2)) + 2.0 * float(np.sum(aa * aa))) / 4.0
    scale = 1.0 / np.sqrt(s1 + s2)
    fbar = f_ov + 2.0 * np.einsum("kd,iakd->ia", t1, g) - np.einsum("kd,idka->ia", t1, g)
    num = 2.0 * float(np.sum(fbar * r1)) + float(np.sum((2.0 * g.transpose(0, 2, 1, 3)
                                                         - g.transpose(0, 2, 3, 1)) * r2))
    return abs(scale * num / omega)


class _Layout:
    """Slots of every molecule's device vectors: two regions of basis V and
    sigmas S (cap each; a restart writes the new basis into the other region),
    the current Ritz vectors with their sigmas (R rows each), and the
    candidates for new basis vectors (corrections, start vectors, seeds)."""

    def __init__(self, nroots, max_space):
        self.nwork = nroots + 4
        self.R = 2 * self.nwork            # a real Ritz vector takes one row, a complex pair two
        self.max_space = max(max_space, 6 * self.nwork)
        self.cap = self.max_space + self.R + 8
        self.VS = [(0, self.cap), (2 * self.cap, 3 * self.cap)]
        self.XB = 4 * self.cap             # X rows at XB, SX rows at XB + R
        self.C = self.XB + 2 * self.R
        self.nC = self.R + self.nwork + 6
        self.nslots = self.C + self.nC


class _Sub:
    """Host state of one molecule's subspace."""

    def __init__(self, o, v, e, nroots, L):
        self.o, self.v = o, v
        self.n1, self.n2 = o * v, o * o * v * v
        self.N = self.n1 + self.n2
        eo, ev = e[:o], e[o:]
        self._d1 = ev[None, :] - eo[:, None]
        self._d2 = (ev[None, None, :, None] + ev[None, None, None, :] - eo[:, None, None, None]
                    - eo[None, :, None, None])
        self.nroots = nroots
        # the space's true dimension: singles plus one coordinate per pair (i a) <= (j b)
        self.dim = self.n1 + self.n1 * (self.n1 + 1) // 2
        # the subspace is expanded along nwork > nroots Ritz pairs: states whose
        # energy drops only once their doubles enter the subspace start high
        # among the Ritz values and would never be targeted otherwise
        self.nwork = min(L.nwork, self.dim)
        self.max_space = min(L.max_space, self.dim)
        self.k = 0            # basis vectors with sigmas: V[0:k], S[0:k]
        self.pending = 0      # basis vectors V[k:k+pending] awaiting their sigmas
        self.G = np.zeros((0, 0))   # G[a, b] = <V_a, S_b>
        self.reg = 0          # region of L.VS holding V and S
        self.Ycur = np.zeros((0, 0))    # this step's Ritz vectors, basis coordinates (k x rows)
        self.Yprev = np.zeros((0, 0))   # the last step's, in the current basis (zero-padded)
        self.rows = 0
        self.seeds = None
        self.omega = np.full(nroots, np.nan)
        self.res = np.full(nroots, np.inf)
        self.root_slot = [None] * nroots
        self.done = False
        self.working = 0      # Ritz values in the working set at the last step

    def V(self, L):
        return L.VS[self.reg][0]

    def S(self, L):
        return L.VS[self.reg][1]

    def unit(self, e1, e2=-1):
        x = np.zeros(self.N)
        x[e1] = 1.0
        if e2 >= 0:
            x[e2] = 1.0
        return x / np.linalg.norm(x)

    def double_seeds(self, count):
        """The ``count`` pair-symmetric unit doubles (i a)(j b) of lowest diagonal."""
        o, v, n1 = self.o, self.v, self.n1
        if n1 < 2 or count <= 0:
            return []
        m, n = np.triu_indices(n1)
        d = self._d2[m // v, n // v, m % v, n % v]
        # the first ``count`` of a stable sort, without sorting all pairs
        if count < len(d):
            cand = np.flatnonzero(d <= np.partition(d, count - 1)[count - 1])
            first = cand[np.argsort(d[cand], kind="stable")][:count]
        else:
            first = np.argsort(d, kind="stable")
        out = []
        for q in first:
            i, a, j, b = m[q] // v, m[q] % v, n[q] // v, n[q] % v
            out.append((n1 + ((i * o + j) * v + a) * v + b, n1 + ((j * o + i) * v + b) * v + a))
        return out

    def next_seeds(self, count):
        """Further start directions in order of the diagonal: unit singles,
        then pair-symmetric unit doubles (for spaces the singles do not fill)."""
        if self.seeds is None:
            o, v, n1 = self.o, self.v, self.n1
            m, n = np.triu_indices(n1)
            i, a, j, b = m // v, m % v, n // v, n % v
            e1 = ((i * o + j) * v + a) * v + b
            e2 = ((j * o + i) * v + b) * v + a
            diag = np.concatenate([self._d1.ravel(), self._d2[i, j, a, b]])
            first = np.concatenate([np.arange(n1), n1 + e1])
            second = np.concatenate([np.full(n1, -1), n1 + e2])
            order = np.argsort(diag, kind="stable")
            self.seeds = list(zip(first[order].tolist(), second[order].tolist()))
            self.seeds.reverse()   # pop from the end
        out = []
        while self.seeds and len(out) < count:
            out.append(self.seeds.pop())
        return out


class _Ops:
    """Batched calls on the device space, with job lists built per molecule
    (subspace k lives in the space of loaded molecule home[k])."""

    def __init__(self, ctx, home):
        self.ctx = ctx
        self.home = np.asarray(home, dtype=np.int32)

    def dots(self, jobs, mode=0):
        jobs = np.array(jobs, dtype=np.int32).reshape(-1, 3)
        jobs[:, 0] = self.home[jobs[:, 0]]
        return self.ctx.lr_dots(jobs, mode)

    def combine(self, jobs):
        """jobs: list of (mol, dst, beta, [(slot, coef), ...])."""
        if not jobs:
            return
        head, beta, ts, tc = [], [], [], []
        for mol, dst, b, terms in jobs:
            head.append((self.home[mol], dst, len(ts), len(terms)))
            beta.append(b)
            for sl, c in terms:
                ts.append(sl)
                tc.append(c)
        self.ctx.lr_combine(np.array(head, dtype=np.int32), np.array(beta), np.array(ts, dtype=np.int32),
                            np.array(tc))

    def combine_blocks(self, blocks):
        """blocks: list of (mol, dst slots [m], beta, term slots [t], coefficients [m, t]):
        dst_q = beta dst_q + sum_j coef[q, j] slot_j, each term read once per 16 rows."""
        heads, betas, sls, cfs = [], [], [], []
        ns = nc = 0
        for mol, dst, beta, slots, coef in blocks:
            m, t = coef.shape
            if m == 0:
                continue
            heads.append((self.home[mol], m, t, ns, ns + m, nc))
            betas.append(beta)
            sls.append(np.asarray(dst, dtype=np.int32))
            sls.append(np.asarray(slots, dtype=np.int32))
            cfs.append(np.asarray(coef, dtype=np.float64).ravel())
            ns += m + t
            nc += m * t
        if heads:
            self.ctx.lr_combine_block(np.array(heads, dtype=np.int32), np.array(betas), np.concatenate(sls),
                                      np.concatenate(cfs))

    def gram(self, blocks):
        """blocks: list of (mol, slots a [na], slots b [nb]) -> list of the
        na x nb matrices <a_p, b_q>, each vector read once per 16 x 64 piece."""
        heads, sls = [], []
        ns = 0
        for mol, a, b in blocks:
            heads.append((self.home[mol], len(a), len(b), ns, ns + len(a)))
            sls.append(np.asarray(a, dtype=np.int32))
            sls.append(np.asarray(b, dtype=np.int32))
            ns += len(a) + len(b)
        if not heads:
            return []
        g = self.ctx.lr_gram(np.array(heads, dtype=np.int32), np.concatenate(sls))
        out, at = [], 0
        for _, na, nb, _, _ in heads:
            out.append(g[at:at + na * nb].reshape(na, nb))
            at += na * nb
        return out


def _accept(M, n0, thresh, room):
    """Gram-Schmidt in the coordinates of the candidates, from their Gram
    matrix M: rows of T give the accepted orthonormal vectors as combinations
    of the candidates, in order; a candidate whose part orthogonal to the
    accepted ones is below ``thresh`` of its original norm n0 is dropped (the
    projection onto the accepted ones applied twice)."""
    m = len(M)
    T = np.zeros((0, m))
    for r in range(m):
        if len(T) >= room:
            break
        c = np.zeros(m)
        c[r] = 1.0
        for _ in range(2):
            c -= T.T @ (T @ (M @ c))
        nrm = np.sqrt(max(float(c @ M @ c), 0.0))
        if nrm > thresh * n0[r]:
            T = np.vstack([T, c / nrm])
    return T


def _orthonormalize(ops, L, subs, cands, thresh):
    """Append the candidates (per molecule: list of (slot, norm before the
    projections)) to the basis as V[k + pending...]: their components along
    the basis are removed by block Gram-Schmidt (twice), then they are
    orthonormalised among themselves in order from their Gram matrix; a
    candidate that keeps less than ``thresh`` of its norm is already held by
    the basis and is dropped."""
    act = [k for k, c in cands.items() if c and subs[k].k + subs[k].pending < min(subs[k].dim, L.cap)]
    if not act:
        return
    for _ in range(2):
        gb = [(k, [sl for sl, _ in cands[k]], subs[k].V(L) + np.arange(subs[k].k + subs[k].pending))
              for k in act if subs[k].k + subs[k].pending > 0]
        if not gb:
            break
        P = ops.gram(gb)
        ops.combine_blocks([(k, cs, 1.0, basis, -p) for (k, cs, basis), p in zip(gb, P)])
    Ms = ops.gram([(k, [sl for sl, _ in cands[k]], [sl for sl, _ in cands[k]]) for k in act])
    blocks = []
    for k, M in zip(act, Ms):
        s = subs[k]
        cs = [sl for sl, _ in cands[k]]
        M = 0.5 * (M + M.T)
        room = min(s.dim, L.cap) - s.k - s.pending
        T = _accept(M, np.array([n0 for _, n0 in cands[k]]), thresh, room)
        if len(T):
            blocks.append((k, s.V(L) + s.k + s.pending + np.arange(len(T)), 0.0, cs, T))
            s.pending += len(T)
    ops.combine_blocks(blocks)


def _collapse(ops, L, subs, which):
    """Restart the subspaces of molecules ``which`` on their current and last
    Ritz vectors; the last step's vectors keep the direction of approach,
    which a restart onto the current ones alone would lose.  Both are
    combinations of the orthonormal basis, so the restart is orthonormalised
    in basis coordinates (Gram-Schmidt applied twice; a vector keeping less
    than 1e-6 of its norm is dropped): V' = V Q and S' = S Q, written into the
    other region, and G' = Q^T G Q."""
    blocks = []
    for k in which:
        s = subs[k]
        kk = s.k
        prev = np.zeros((kk, s.Yprev.shape[1]))
        prev[:len(s.Yprev)] = s.Yprev
        W = np.concatenate([s.Ycur, prev], axis=1)
        Q = np.zeros((kk, 0))
        for j in range(W.shape[1]):
            c = W[:, j].copy()
            n0 = np.linalg.norm(c)
            for _ in range(2):
                c -= Q @ (Q.T @ c)
            nr = np.linalg.norm(c)
            if nr > 1e-6 * n0 and Q.shape[1] < kk:
                Q = np.concatenate([Q, (c / nr)[:, None]], axis=1)
        v1, s1 = L.VS[1 - s.reg]
        blocks.append((k, v1 + np.arange(Q.shape[1]), 0.0, s.V(L) + np.arange(kk), Q.T))
        blocks.append((k, s1 + np.arange(Q.shape[1]), 0.0, s.S(L) + np.arange(kk), Q.T))
        s.G = Q.T @ s.G @ Q
        s.Ycur = Q.T @ s.Ycur
        s.reg = 1 - s.reg
        s.k = Q.shape[1]
    ops.combine_blocks(blocks)


def _absorb(ops, L, subs, act):
    """Extend G by the rows and columns of the vectors whose sigmas just arrived."""
    gb = []
    for k in act:
        s = subs[k]
        k0, k1 = s.k, s.k + s.pending
        gb.append((k, s.V(L) + np.arange(k1), s.S(L) + np.arange(k0, k1)))      # new columns, all rows
        gb.append((k, s.V(L) + np.arange(k0, k1), s.S(L) + np.arange(k0)))      # new rows, old columns
    g = ops.gram(gb)
    for q, k in enumerate(act):
        s = subs[k]
        k0, k1 = s.k, s.k + s.pending
        G = np.zeros((k1, k1))
        G[:k0, :k0] = s.G
        G[:, k0:] = g[2 * q]
        G[k0:, :k0] = g[2 * q + 1]
        s.G = G
        s.k, s.pending = k1, 0


def _ritz_all(subs, act):
    """_ritz of the molecules ``act``, the eigenproblems of equal dimension
    solved as one stacked call."""
    by_k = {}
    for k in act:
        by_k.setdefault(subs[k].k, []).append(k)
    out = {}
    for _, ks in by_k.items():
        lam, Y = np.linalg.eig(np.stack([subs[k].G for k in ks]))
        for q, k in enumerate(ks):
            out[k] = _ritz(subs[k], lam[q], Y[q])
    return out


def _ritz(s, lam, Y):
    """Rayleigh-Ritz of one subspace from the eigenpairs (lam, Y) of
    G = V^T J V: the working Ritz values, lowest real parts first (one entry
    per real value and per complex pair), and the subspace coefficients of
    their vectors (a real one: one row; a complex pair: its real and
    imaginary parts)."""
    keep = lam.imag >= -1e-12 * np.maximum(np.abs(lam.real), 1.0)
    lam, Y = lam[keep], Y[:, keep]
    order = np.argsort(lam.real, kind="stable")
    lam, Y = lam[order], Y[:, order]
    isreal = np.abs(lam.imag) <= 1e-10 * np.maximum(np.abs(lam.real), 1.0)
    rows, entries = [], []
    for r in range(min(s.nwork, len(lam))):
        y = Y[:, r] / np.linalg.norm(Y[:, r])
        entries.append((lam[r], bool(isreal[r]), len(rows)))
        if isreal[r]:
            rows.append(y.real / np.linalg.norm(y.real))
        else:
            rows.extend([y.real, y.imag])
    return entries, rows


def excitation_energies(ctx, dims, nroots=1, eps=1e-6, maxiter=100, h=1e-5, max_space=48, log=None,
                        gammas=None, replicas=None):
    """Lowest ``nroots`` singlet LR-CCSD excitation energies (eV) of every
    loaded molecule, after a converged solve of ``ctx``'s batch.

    dims: (n_occ, n_virt) of the loaded molecules in load order.
    gammas: the molecules' site matrices gamma; when given, the Davidson starts
    from the lowest CIS singlets (the reference's default 'cis' guess),
    otherwise from unit singles (its 'unt' guess).
    Returns (omega [nmol, nroots], residual max-norms [nmol, nroots],
    iterations [nmol], converged [nmol], average excitation levels
    [nmol, nroots], reference weights r0 [nmol, nroots] -- NaN without
    ``gammas``).  ``log(iteration, active)`` is called after every Davidson
    iteration when given.

    replicas: for a batch that holds several copies of each molecule (loaded
    from the same inputs and solved alike, so with the same amplitudes), the
    batch indices of each molecule's copies, one list per entry of ``dims``;
    the first holds the molecule's subspace, and one Jacobian product call
    applies J to as many of its vectors as it has copies (the products are
    those of the single molecule, bit for bit)."""
    # the host's matrices are small (projected problems of a few dozen rows,
    # CIS matrices of a few hundred): several BLAS threads only compete
    with threadpool_limits(1):
        return _excitation_energies(ctx, dims, nroots, eps, maxiter, h, max_space, log, gammas, replicas)


def _excitation_energies(ctx, dims, nroots, eps, maxiter, h, max_space, log, gammas, replicas):
    L = _Layout(nroots, max_space)
    home = [r[0] for r in replicas] if replicas is not None else list(range(len(dims)))
    if replicas is not None:
        counts = np.zeros(ctx.nmol, dtype=np.int32)
        counts[home] = L.nslots
        ctx.lr_space_counts(counts)
    else:
        ctx.lr_space(L.nslots)
    ops = _Ops(ctx, home)
    subs = []
    cands = {}
    for k, (o, v) in enumerate(dims):
        C, Fmo, _ = ctx.batch_orbitals(home[k], o + v)
        e = np.diag(Fmo).copy()
        s = _Sub(o, v, e, nroots, L)
        subs.append(s)
        # more start vectors than roots: a Davidson subspace keeps the
        # point-group symmetry of its start, so every symmetry among the low
        # states needs one.  Either the lowest CIS states or unit singles on
        # the lowest orbital-energy differences; then the lowest
        # pair-symmetric doubles: states of mainly double excitation
        # character can lie far below their diagonal, and reached only
        # through the singles' corrections they surface late.
        count = min(s.n1, nroots + 6)
        starts = []
        if gammas is not None:
            U = cis_singlets(C, np.asarray(gammas[k], dtype=np.float64), e, o, count)[1]
            for q in range(count):
                x = np.zeros(s.N)
                x[:s.n1] = U[:, q] / np.linalg.norm(U[:, q])
                starts.append(x)
        else:
            for idx in np.argsort(s._d1.ravel(), kind="stable")[:count]:
                starts.append(s.unit(int(idx)))
        starts.extend(s.unit(e1, e2) for e1, e2 in s.double_seeds(2))
        cands[k] = []
        for q, x in enumerate(starts):
            ctx.lr_put(home[k], L.C + q, x)
            cands[k].append((L.C + q, 1.0))
    _orthonormalize(ops, L, subs, cands, 1e-3)
    its = np.zeros(len(subs), dtype=int)
    for it in range(1, maxiter + 1):
        act = [k for k, s in enumerate(subs) if not s.done and s.pending > 0]
        for k, s in enumerate(subs):
            if not s.done and s.pending == 0:
                s.done = True
        if not act:
            break
        # sigmas of the pending vectors: one batched product per column, or
        # with replicas, one per len(replicas[k]) columns
        if replicas is None:
            for c in range(max(subs[k].pending for k in act)):
                ins = np.full(len(subs), -1, dtype=np.int32)
                outs = np.zeros(len(subs), dtype=np.int32)
                for k in act:
                    s = subs[k]
                    if c < s.pending:
                        ins[k] = s.V(L) + s.k + c
                        outs[k] = s.S(L) + s.k + c
                ctx.lr_sigma_slots(ins, outs, h)
        else:
            rounds = max(-(-subs[k].pending // len(replicas[k])) for k in act)
            for r in range(rounds):
                src = np.zeros(ctx.nmol, dtype=np.int32)
                ins = np.full(ctx.nmol, -1, dtype=np.int32)
                outs = np.zeros(ctx.nmol, dtype=np.int32)
                for k in act:
                    s = subs[k]
                    for j, b in enumerate(replicas[k]):
                        c = r * len(replicas[k]) + j
                        if c < s.pending:
                            src[b] = home[k]
                            ins[b] = s.V(L) + s.k + c
                            outs[b] = s.S(L) + s.k + c
                ctx.lr_sigma_from(src, ins, outs, h)
        _absorb(ops, L, subs, act)
        # Ritz vectors and their sigmas into the Ritz buffer
        ritz = {}
        blocks = []
        xb = L.XB
        found = _ritz_all(subs, act)
        for k in act:
            s = subs[k]
            entries, rows = found[k]
            ritz[k] = entries
            s.rows = len(rows)
            Y = np.array(rows).reshape(-1, s.k)
            s.Ycur = Y.T
            blocks.append((k, xb + np.arange(len(Y)), 0.0, s.V(L) + np.arange(s.k), Y))
            blocks.append((k, xb + L.R + np.arange(len(Y)), 0.0, s.S(L) + np.arange(s.k), Y))
        ops.combine_blocks(blocks)
        # residuals J x - lam x into the candidate slots, and their max-norms
        cj, mj, where = [], [], {}
        for k in act:
            s = subs[k]
            q = 0
            where[k] = []
            for lam, real, c in ritz[k]:
                X, SX = xb + c, xb + L.R + c
                if real:
                    cj.append((k, L.C + q, 0.0, [(SX, 1.0), (X, -lam.real)]))
                    mj.append((k, L.C + q, -1))
                    where[k].append((q, 1))
                    q += 1
                else:
                    a, b = lam.real, lam.imag
                    cj.append((k, L.C + q, 0.0, [(SX, 1.0), (X, -a), (X + 1, b)]))
                    cj.append((k, L.C + q + 1, 0.0, [(SX + 1, 1.0), (X + 1, -a), (X, -b)]))
                    mj.append((k, L.C + q, L.C + q + 1))
                    where[k].append((q, 2))
                    q += 2
        ops.combine(cj)
        resn = ops.dots(mj, mode=1)
        # convergence, reported roots, corrections to keep
        at = 0
        corr = {}
        pj, shifts = [], []
        collapse = []
        for k in act:
            s = subs[k]
            its[k] = it
            reals, worst_pair, keep = [], 0.0, []
            for (lam, real, c), (q, width) in zip(ritz[k], where[k]):
                res = resn[at]
                at += 1
                if real:
                    reals.append((lam.real, res, c))
                elif len(reals) < s.nroots:
                    worst_pair = max(worst_pair, res)
                if res >= eps:
                    for w in range(width):
                        keep.append(L.C + q + w)
                        pj.append((k, L.C + q + w))
                        shifts.append(lam.real)
            s.working = len(ritz[k])
            for r, (val, res, c) in enumerate(reals[:s.nroots]):
                s.omega[r] = val
                s.res[r] = res
                s.root_slot[r] = L.XB + c
            nr = min(s.nroots, s.dim)
            if len(reals) >= nr and all(res < eps for _, res, _ in reals[:nr]) and worst_pair < 1e-3:
                s.done = True
                continue
            if s.k >= s.dim:   # the subspace spans the space: Ritz values are exact
                s.done = True
                continue
            corr[k] = keep
            if s.k + len(keep) > s.max_space:
                collapse.append(k)
        ops_pre = [(home[k], sl) for k, sl in pj if k in corr]
        ctx.lr_precondition(np.array(ops_pre, dtype=np.int32).reshape(-1, 2),
                            np.array([sh for (k, _), sh in zip(pj, shifts) if k in corr]), 1e-2)
        live = list(corr)
        n0 = ops.dots([(k, sl, sl) for k in live for sl in corr[k]])
        cands, at = {}, 0
        for k in live:
            cands[k] = [(sl, np.sqrt(max(n0[at + q], 0.0))) for q, sl in enumerate(corr[k])]
            at += len(corr[k])
        if collapse:
            _collapse(ops, L, subs, collapse)
        for k in act:
            subs[k].Yprev = subs[k].Ycur
        _orthonormalize(ops, L, subs, cands, 1e-3)
        # fewer Ritz values than working roots: widen the subspace along the
        # next diagonal directions
        for _ in range(8):
            seeds = {}
            for k in live:
                s = subs[k]
                need = s.nwork - s.working - s.pending
                room = min(s.dim, L.cap) - s.k - s.pending
                if need > 0 and room > 0:
                    got = s.next_seeds(min(need, room))
                    if got:
                        seeds[k] = []
                        for q, (e1, e2) in enumerate(got):
                            ctx.lr_put(home[k], L.C + q, s.unit(e1, e2))
                            seeds[k].append((L.C + q, 1.0))
            if not seeds:
                break
            _orthonormalize(ops, L, subs, seeds, 1e-3)
        if log is not None:
            log(it, sum(not s.done for s in subs))
    omega = np.array([s.omega for s in subs])
    res = np.array([s.res for s in subs])
    conv = np.array([bool(np.all(s.res[:min(nroots, s.dim)] < eps)) or s.k >= s.dim for s in subs])
    ael = np.full((len(subs), nroots), np.nan)
    r0 = np.full((len(subs), nroots), np.nan)
    for k, s in enumerate(subs):
        g = f_ov = t1 = None
        if gammas is not None:
            C, Fmo, _ = ctx.batch_orbitals(home[k], s.o + s.v)
            g = ovov_integrals(C, np.asarray(gammas[k], dtype=np.float64), s.o)
            f_ov = Fmo[:s.o, s.o:]
            t1, _ = ctx.batch_amplitudes(home[k], s.o, s.v, doubles=False)
        for r, sl in enumerate(s.root_slot):
            if sl is not None:
                x = ctx.lr_get(home[k], sl, s.N)
                x1, x2 = x[:s.n1].reshape(s.o, s.v), x[s.n1:].reshape(s.o, s.o, s.v, s.v)
                ael[k, r] = average_excitation_level(x1, x2)
                if g is not None:
                    r0[k, r] = reference_weight(g, f_ov, t1, s.omega[r], x1, x2)
    return omega, res, its, conv, ael, r0
