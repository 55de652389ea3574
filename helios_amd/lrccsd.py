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
J = dOmega/dt at the converged amplitudes, applied on the GPU for every
molecule of the batch at once (``hcc_batch_lr_sigma``: central differences
of the FP64 residual), with a block Davidson solve on the host for the
``nroots`` lowest roots of each molecule: Rayleigh-Ritz on an orthonormal
subspace, residual corrections preconditioned by the orbital-energy
differences, subspace collapse onto the Ritz vectors when it grows too
large.  Trial vectors start as unit singles on the lowest orbital-energy
differences (the reference's ``unt`` guess).  The difference step h = 1e-5
keeps sigma within about 1e-9 of J r (the truncation error is O(h^2) times
the third derivative of a quartic polynomial; the FP64 rounding error is
about 1e-16 |Omega| / h).  Triplets (spin = 1) need the
spin-orbital Jacobian and are not provided.
"""
import numpy as np

from helios_amd import capi


class _Davidson:
    """Subspace of one molecule: orthonormal basis rows Vm[:k], sigma rows Sm[:k]."""

    def __init__(self, o, v, e, nroots, max_space):
        self.o, self.v = o, v
        self.n1, self.n2 = o * v, o * o * v * v
        N = self.n1 + self.n2
        eo, ev = e[:o], e[o:]
        d1 = ev[None, :] - eo[:, None]
        d2 = (ev[None, None, :, None] + ev[None, None, None, :] - eo[:, None, None, None]
              - eo[None, :, None, None])
        self.D = np.concatenate([d1.ravel(), d2.ravel()])
        self.nroots = nroots
        # the subspace is expanded along nwork > nroots Ritz pairs: states whose
        # energy drops only once their doubles enter the subspace start high
        # among the Ritz values and would never be targeted otherwise
        self.nwork = nroots + 4
        # the space's true dimension: singles plus one coordinate per pair (i a) <= (j b)
        self.dim = self.n1 + self.n1 * (self.n1 + 1) // 2
        self.nwork = min(self.nwork, self.dim)
        self.max_space = min(max(max_space, 6 * self.nwork), self.dim)
        cap = self.max_space + 2 * self.nwork + 8
        self.Vm = np.zeros((cap, N))
        self.Sm = np.zeros((cap, N))
        self.k = 0
        self.G = np.zeros((0, 0))   # G[k, l] = <v_k, J v_l>, grown with the subspace
        self.new = []
        # unit singles on the lowest orbital-energy differences, more than the
        # roots: a Davidson subspace keeps the point-group symmetry of its
        # start, so every symmetry among the low states needs a start vector
        for idx in np.argsort(d1.ravel(), kind="stable")[:min(self.n1, nroots + 6)]:
            x = np.zeros(N)
            x[idx] = 1.0
            self.new.append(x)
        # further start directions, in order of the diagonal: unit singles, then
        # pair-symmetric unit doubles (for spaces the singles do not fill);
        # ordered on first use
        self._d1, self._d2 = d1, d2
        self.seeds = None
        # the lowest pair-symmetric doubles start too: states of mainly double
        # excitation character can lie far below their diagonal, and reached
        # only through the singles' corrections they surface late
        for e1, e2 in self._double_seeds(2):
            x = np.zeros(N)
            x[e1] = x[e2] = 1.0
            self.new.append(x / np.linalg.norm(x))
        self.prev = []     # last step's Ritz vectors (x, J x), kept through a collapse
        self.omega = np.full(nroots, np.nan)
        self.res = np.full(nroots, np.inf)
        self.done = False

    def _double_seeds(self, count):
        """The ``count`` pair-symmetric unit doubles (i a)(j b) of lowest diagonal."""
        o, v, n1 = self.o, self.v, self.n1
        if self.n1 < 2 or count <= 0:
            return []
        m, n = np.triu_indices(n1)
        d = self._d2[m // v, n // v, m % v, n % v]
        out = []
        for q in np.argsort(d, kind="stable")[:count]:
            i, a, j, b = m[q] // v, m[q] % v, n[q] // v, n[q] % v
            out.append((n1 + ((i * o + j) * v + a) * v + b, n1 + ((j * o + i) * v + b) * v + a))
        return out

    def _seed_order(self):
        o, v, n1 = self.o, self.v, self.n1
        m, n = np.triu_indices(n1)                  # pairs (i a) = m <= n = (j b)
        i, a, j, b = m // v, m % v, n // v, n % v
        e1 = ((i * o + j) * v + a) * v + b
        e2 = ((j * o + i) * v + b) * v + a
        diag = np.concatenate([self._d1.ravel(), self._d2[i, j, a, b]])
        first = np.concatenate([np.arange(n1), n1 + e1])
        second = np.concatenate([np.full(n1, -1), n1 + e2])
        order = np.argsort(diag, kind="stable")
        return list(zip(first[order].tolist(), second[order].tolist()))

    def _seed(self):
        """Next start direction not yet in the subspace (None when exhausted)."""
        if self.seeds is None:
            self.seeds = self._seed_order()
        while self.seeds:
            e1, e2 = self.seeds.pop(0)
            x = np.zeros(self.n1 + self.n2)
            x[e1] = 1.0
            if e2 >= 0:
                x[e2] = 1.0
            n0 = np.linalg.norm(x)
            x = self._orth(x)
            nrm = np.linalg.norm(x)
            if nrm > 1e-3 * n0:
                return x / nrm
        return None

    def _orth(self, x):
        V = self.Vm[:self.k]
        for _ in range(2):
            x = x - (V @ x) @ V
            for b in self.new:
                x = x - (b @ x) * b
        return x

    def absorb(self, sig):
        """Add the pending trial vectors and their sigma vectors."""
        k0, m = self.k, len(self.new)
        for q in range(m):
            self.Vm[k0 + q] = self.new[q]
            self.Sm[k0 + q] = sig[q]
        self.k = k0 + m
        self.new = []
        V, S = self.Vm[:self.k], self.Sm[:self.k]
        G = np.zeros((self.k, self.k))
        G[:k0, :k0] = self.G
        G[:, k0:] = V @ S[k0:].T
        G[k0:, :k0] = V[k0:] @ S[:k0].T
        self.G = G

    def step(self, eps):
        V, S = self.Vm[:self.k], self.Sm[:self.k]
        lam, Y = np.linalg.eig(self.G)   # right eigenvectors of V^T J V
        # One entry per real Ritz value and per complex pair (J is not
        # symmetric), lowest real parts first.  Complex pairs are not
        # excitation energies; they stay in the working set (expanded along
        # the real and imaginary parts of their corrections) so that a pair
        # close to the real axis is resolved rather than read as a root.
        keep = lam.imag >= -1e-12 * np.maximum(np.abs(lam.real), 1.0)
        lam, Y = lam[keep], Y[:, keep]
        order = np.argsort(lam.real, kind="stable")
        lam, Y = lam[order], Y[:, order]
        isreal = np.abs(lam.imag) <= 1e-10 * np.maximum(np.abs(lam.real), 1.0)
        k = min(self.nwork, len(lam))
        corrections = []
        reals = []          # (value, residual, x, J x) of the real Ritz values in the working set
        pairs = []          # real and imaginary parts of the complex Ritz pairs' (x, J x)
        worst_pair = 0.0    # largest residual of a complex pair below the highest reported root
        for r in range(k):
            y = Y[:, r] / np.linalg.norm(Y[:, r])
            if isreal[r]:
                y = y.real / np.linalg.norm(y.real)
                x = y @ V
                q = y @ S - lam[r].real * x
                res = np.abs(q).max()
                reals.append((lam[r].real, res, x, y @ S))
                parts = [q] if res >= eps else []
            else:
                x = y @ V
                sx = y @ S
                q = sx - lam[r] * x
                res = np.abs(q).max()
                if len(reals) < self.nroots:
                    worst_pair = max(worst_pair, res)
                pairs.extend([(x.real, sx.real), (x.imag, sx.imag)])
                parts = [q.real, q.imag] if res >= eps else []
            den = lam[r].real - self.D
            den = np.where(np.abs(den) < 1e-2, np.copysign(1e-2, den), den)
            corrections.extend(c / den for c in parts)
        for r, (val, res, _, _) in enumerate(reals[:self.nroots]):
            self.omega[r] = val
            self.res[r] = res
        ritz = [(x, sx) for _, _, x, sx in reals] + pairs
        nr = min(self.nroots, self.dim)
        if len(reals) >= nr and all(res < eps for _, res, _, _ in reals[:nr]) and worst_pair < 1e-3:
            self.done = True
            return
        if self.k >= self.dim:   # the subspace spans the space: Ritz values are exact
            self.done = True
            return
        if self.k + len(corrections) > self.max_space:
            # collapse onto the orthonormalised Ritz vectors of this step and
            # of the last one (their sigmas follow linearly): the last step's
            # vectors keep the direction of approach, which a restart onto
            # the current ones alone would lose
            basis = ritz + self.prev
            self.k = 0
            for x, sx in basis:
                n0 = np.linalg.norm(x)
                for _ in range(2):
                    c = self.Vm[:self.k] @ x
                    x = x - c @ self.Vm[:self.k]
                    sx = sx - c @ self.Sm[:self.k]
                nrm = np.linalg.norm(x)
                if nrm > 1e-6 * n0:
                    self.Vm[self.k] = x / nrm
                    self.Sm[self.k] = sx / nrm
                    self.k += 1
            self.G = self.Vm[:self.k] @ self.Sm[:self.k].T
        self.prev = ritz
        for c in corrections:
            n0 = np.linalg.norm(c)
            c = self._orth(c)
            nrm = np.linalg.norm(c)
            # keep a correction unless the subspace already contains it
            if nrm > 1e-3 * n0 and self.k + len(self.new) < self.dim:
                self.new.append(c / nrm)
        # fewer real Ritz values than working roots: widen the subspace along
        # the next diagonal directions
        while k + len(self.new) < self.nwork and self.k + len(self.new) < self.dim:
            x = self._seed()
            if x is None:
                break
            self.new.append(x)
        if not self.new:
            self.done = True


def excitation_energies(ctx, dims, nroots=1, eps=1e-6, maxiter=100, h=1e-5, max_space=48, log=None):
    """Lowest ``nroots`` singlet LR-CCSD excitation energies (eV) of every
    loaded molecule, after a converged solve of ``ctx``'s batch.

    dims: (n_occ, n_virt) of the loaded molecules in load order.
    Returns (omega [nmol, nroots], residual max-norms [nmol, nroots],
    iterations [nmol], converged [nmol]).  ``log(iteration, active)`` is called
    after every Davidson iteration when given."""
    subs = []
    for k, (o, v) in enumerate(dims):
        _, Fmo, _ = ctx.batch_orbitals(k, o + v)
        subs.append(_Davidson(o, v, np.diag(Fmo).copy(), nroots, max_space))
    n1 = [s.n1 for s in subs]
    n2 = [s.n2 for s in subs]
    o1 = np.concatenate([[0], np.cumsum(n1)])
    o2 = np.concatenate([[0], np.cumsum(n2)])
    its = np.zeros(len(subs), dtype=int)
    for it in range(1, maxiter + 1):
        width = max((len(s.new) for s in subs if not s.done), default=0)
        if width == 0:
            break
        sig = [[] for _ in subs]
        for c in range(width):
            r1 = np.zeros(o1[-1])
            r2 = np.zeros(o2[-1])
            for k, s in enumerate(subs):
                if not s.done and c < len(s.new):
                    r1[o1[k]:o1[k + 1]] = s.new[c][:s.n1]
                    r2[o2[k]:o2[k + 1]] = s.new[c][s.n1:]
            s1, s2 = ctx.batch_lr_sigma(r1, r2, h)
            for k, s in enumerate(subs):
                if not s.done and c < len(s.new):
                    sig[k].append(np.concatenate([s1[o1[k]:o1[k + 1]], s2[o2[k]:o2[k + 1]]]))
        for k, s in enumerate(subs):
            if s.done:
                continue
            s.absorb(sig[k])
            s.step(eps)
            its[k] = it
        if log is not None:
            log(it, sum(not s.done for s in subs))
    omega = np.array([s.omega for s in subs])
    res = np.array([s.res for s in subs])
    conv = np.array([bool(np.all(s.res[:min(nroots, s.dim)] < eps)) or s.k >= s.dim for s in subs])
    return omega, res, its, conv
