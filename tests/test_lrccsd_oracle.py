"""Singlet LR-CCSD oracle and host solver (CPU).

* Pin: for two electrons EOM-CCSD is exact, so every eigenvalue of the
  closed-shell singlet Jacobian (dimension ov + ov(ov+1)/2 = n(n+1)/2 - 1)
  equals an excitation energy of the two-electron singlet full CI, which is
  written independently in the MO product basis.
* The block Davidson of helios_amd.lrccsd, driven by the oracle's sigma
  (the same central difference the GPU computes), finds the three lowest
  real eigenvalues of the dense Jacobian, including the doubly excited
  ("dark") lowest state of the polyenes, from unit singles and from the CIS
  start vectors.
* The CIS start vectors are eigenpairs of the spin-orbital CIS matrix; the
  average excitation level equals the reference's spin-orbital sum.
* Driven with copies of each molecule (several Jacobian products per call),
  the Davidson gives the single-copy roots, residuals and iteration counts.
"""
import numpy as np
import pytest

from helios_amd import lrccsd, molecules as M
from oracle import ccsd, lrccsd as olr, ppp, scf
from tests._molprep import prepare


def _two_electrons(mol):
    G, core, hk = ppp.build(mol.coords, mol.bonds, alpha=mol.alpha, U=mol.U, kind="ohno-klopman")
    r = scf.rhf(core, G, hk, 1)
    F, eref = scf.mo_fock(r["C"], core, G, 1)
    return G, core, r["C"], F, eref


@pytest.mark.parametrize("mol", [M.ethylene(), M.polyene(4), M.random_fragment(6, seed=3)], ids=lambda m: m.name)
def test_two_electron_jacobian_spectrum_is_fci(mol):
    G, core, C, F, eref = _two_electrons(mol)
    r = ccsd.solve(C, F, G, 1, eref, eps=1e-11, diis_scale=True)
    assert r["status"] == 1
    lam, ncomplex = olr.excitation_energies(C, F, G, 1, r["t1"], r["t2"])
    fci = olr.singlets_two_electrons(C, core, G)
    n = C.shape[0]
    assert ncomplex == 0 and len(lam) == len(fci) == n * (n + 1) // 2 - 1
    assert np.abs(lam - fci).max() < 1e-8


class _OracleCtx:
    """The calls helios_amd.lrccsd makes, answered by the oracle: the device
    vector space as numpy arrays with the C ABI's semantics, sigma by the
    oracle's central difference."""

    def __init__(self, items):
        self.items = items
        self.space = None
        self.nmol = len(items)

    def batch_orbitals(self, k, n):
        p = self.items[k][0]
        return p["C"], p["F"], p["eref"]

    def batch_amplitudes(self, k, o, v, doubles=True):
        return self.items[k][2], (self.items[k][3] if doubles else None)

    def batch_lr_sigma(self, r1, r2, h):
        s1, s2 = np.zeros_like(r1), np.zeros_like(r2)
        a1 = a2 = 0
        for p, o, t1, t2 in self.items:
            v = p["n"] - o
            n1, n2 = o * v, o * o * v * v
            q1, q2 = olr.sigma(p["C"], p["F"], p["G"], o, t1, t2, r1[a1:a1 + n1].reshape(o, v),
                               r2[a2:a2 + n2].reshape(o, o, v, v), h)
            s1[a1:a1 + n1], s2[a2:a2 + n2] = q1.ravel(), q2.ravel()
            a1, a2 = a1 + n1, a2 + n2
        return s1, s2

    def lr_space(self, nslots):
        self.space = []
        for p, o, _, _ in self.items:
            v = p["n"] - o
            self.space.append(np.zeros((nslots, o * v + o * o * v * v)))

    def lr_space_counts(self, counts):
        self.space = []
        for (p, o, _, _), c in zip(self.items, counts):
            v = p["n"] - o
            self.space.append(np.zeros((c, o * v + o * o * v * v)) if c > 0 else None)

    def lr_sigma_from(self, src, in_slot, out_slot, h):
        writes = []
        for k, (m, a, b) in enumerate(zip(src, in_slot, out_slot)):
            if a < 0:
                continue
            p, o, t1, t2 = self.items[k]          # molecule k's own amplitudes (a copy of m's)
            v = p["n"] - o
            n1 = o * v
            x = self.space[m][a]
            q1, q2 = olr.sigma(p["C"], p["F"], p["G"], o, t1, t2, x[:n1].reshape(o, v),
                               x[n1:].reshape(o, o, v, v), h)
            writes.append((m, b, np.concatenate([q1.ravel(), q2.ravel()])))
        assert len({(m, b) for m, b, _ in writes}) == len(writes)
        for m, b, y in writes:
            self.space[m][b] = y

    def lr_put(self, mol, slot, x):
        self.space[mol][slot] = x

    def lr_get(self, mol, slot, n):
        return self.space[mol][slot].copy()

    def lr_dots(self, jobs, mode=0):
        jobs = np.asarray(jobs, dtype=np.int64).reshape(-1, 3)
        out = np.zeros(len(jobs))
        for j, (m, a, b) in enumerate(jobs):
            x = self.space[m][a]
            if mode == 0:
                out[j] = x @ self.space[m][b]
            else:
                y = self.space[m][b] if b >= 0 else 0.0 * x
                out[j] = np.sqrt(x * x + y * y).max()
        return out

    def lr_combine(self, jobs, beta, term_slot, term_coef):
        for (m, d, f, c), bt in zip(np.asarray(jobs).reshape(-1, 4), beta):
            assert d not in term_slot[f:f + c]
            acc = bt * self.space[m][d] if bt != 0.0 else np.zeros_like(self.space[m][d])
            for t in range(f, f + c):
                acc = acc + term_coef[t] * self.space[m][term_slot[t]]
            self.space[m][d] = acc

    def lr_gram(self, jobs, slots):
        out = []
        for m, na, nb, a, b in np.asarray(jobs, dtype=np.int64).reshape(-1, 5):
            X = self.space[m][slots[a:a + na]]
            Y = self.space[m][slots[b:b + nb]]
            out.append((X @ Y.T).ravel())
        return np.concatenate(out) if out else np.zeros(0)

    def lr_combine_block(self, jobs, beta, slots, coef):
        for (m, mm, t, d, s, c), bt in zip(np.asarray(jobs, dtype=np.int64).reshape(-1, 6), beta):
            dst, src = list(slots[d:d + mm]), list(slots[s:s + t])
            assert len(set(dst)) == mm and not set(dst) & set(src)
            C = np.asarray(coef[c:c + mm * t]).reshape(mm, t)
            acc = C @ self.space[m][src] if t else np.zeros((mm, self.space[m].shape[1]))
            if bt != 0.0:
                acc = acc + bt * self.space[m][dst]
            self.space[m][dst] = acc

    def lr_precondition(self, jobs, shift, floor):
        for (m, sl), w in zip(np.asarray(jobs).reshape(-1, 2), shift):
            p, o = self.items[m][0], self.items[m][1]
            e = np.diag(p["F"])
            d1 = e[o:][None, :] - e[:o][:, None]
            d2 = e[o:][None, None, :, None] + e[o:][None, None, None, :] - e[:o][:, None, None, None] \
                - e[:o][None, :, None, None]
            den = w - np.concatenate([d1.ravel(), d2.ravel()])
            den = np.where(np.abs(den) < floor, np.copysign(floor, den), den)
            self.space[m][sl] /= den

    def lr_sigma_slots(self, in_slot, out_slot, h):
        for m, (a, b) in enumerate(zip(in_slot, out_slot)):
            if a < 0:
                continue
            p, o, t1, t2 = self.items[m]
            v = p["n"] - o
            n1 = o * v
            x = self.space[m][a]
            q1, q2 = olr.sigma(p["C"], p["F"], p["G"], o, t1, t2, x[:n1].reshape(o, v),
                               x[n1:].reshape(o, o, v, v), h)
            self.space[m][b] = np.concatenate([q1.ravel(), q2.ravel()])


@pytest.mark.parametrize("guess", ["unt", "cis"])
def test_davidson_finds_the_lowest_roots(guess):
    items = []
    for mol in [M.octatetraene(), M.polyene(6), M.random_fragment(10, seed=4)]:
        p = prepare(mol)
        r = ccsd.solve(p["C"], p["F"], p["G"], mol.nocc, p["eref"], eps=1e-11, diis_scale=True)
        items.append((p, mol.nocc, r["t1"], r["t2"]))
    gammas = [p["G"] for p, _, _, _ in items] if guess == "cis" else None
    om, res, its, conv, ael, r0 = lrccsd.excitation_energies(
        _OracleCtx(items), [(o, p["n"] - o) for p, o, _, _ in items], nroots=3, gammas=gammas)
    for (p, o, t1, t2), w, c, lv, q in zip(items, om, conv, ael, r0):
        lam, _, vec = olr.excitation_energies(p["C"], p["F"], p["G"], o, t1, t2, vectors=True)
        print(p["n"], w, lam[:3], "AEL", lv, "r0", q)
        assert c
        assert np.abs(w - lam[:3]).max() < 2e-6
        assert np.all((lv >= 1.0) & (lv <= 2.0))
        if guess == "cis":
            want = [olr.reference_weight(p["C"], p["F"], p["G"], o, t1, lam[r], *vec[r]) for r in range(3)]
            assert np.abs(q - want).max() < 1e-5
        else:
            assert np.isnan(q).all()


def test_davidson_with_copies_equals_single_copies():
    """The driver's bookkeeping for copies of a molecule (several Jacobian
    products per call, read from and written to the first copy's slots) gives
    the roots, residuals and iteration counts of the single-copy driver."""
    items = []
    for mol in [M.octatetraene(), M.random_fragment(10, seed=4)]:
        p = prepare(mol)
        r = ccsd.solve(p["C"], p["F"], p["G"], mol.nocc, p["eref"], eps=1e-11, diis_scale=True)
        items.append((p, mol.nocc, r["t1"], r["t2"]))
    dims = [(o, p["n"] - o) for p, o, _, _ in items]
    ref = lrccsd.excitation_energies(_OracleCtx(items), dims, nroots=2)
    copies = 3
    got = lrccsd.excitation_energies(_OracleCtx([it for it in items for _ in range(copies)]), dims, nroots=2,
                                     replicas=[[copies * k + j for j in range(copies)] for k in range(len(items))])
    for a, b in zip(ref[:4], got[:4]):
        assert np.array_equal(a, b)


def test_cis_singlets_are_spin_orbital_eigenpairs():
    """Each singlet CIS pair (value, vector) is an eigenpair of the
    spin-orbital CIS matrix with equal alpha and beta components; a triplet
    coupling (no Coulomb term), one Coulomb term instead of two, or a missing
    exchange term fails (residuals 0.7-4.8 eV on octatetraene)."""
    for mol in [M.octatetraene(), M.random_fragment(10, seed=4)]:
        p = prepare(mol)
        o, n = mol.nocc, p["n"]
        v = n - o
        w, U = lrccsd.cis_singlets(p["C"], p["G"], np.diag(p["F"]), o)
        A = olr.cis_spin_orbital(p["C"], p["F"], p["G"], o)
        for k in range(o * v):
            r = np.zeros((2 * o, 2 * v))
            r[0::2, 0::2] = r[1::2, 1::2] = U[:, k].reshape(o, v)
            x = r.ravel() / np.linalg.norm(r)
            assert np.abs(A @ x - w[k] * x).max() < 1e-10
        assert np.all(np.diff(w) >= 0)
        # the lowest few alone (the Davidson's start vectors): same values, eigenvectors of A
        ws, Us = lrccsd.cis_singlets(p["C"], p["G"], np.diag(p["F"]), o, 5)
        assert ws.shape == (5,) and np.abs(ws - w[:5]).max() < 1e-10
        for k in range(5):
            r = np.zeros((2 * o, 2 * v))
            r[0::2, 0::2] = r[1::2, 1::2] = Us[:, k].reshape(o, v)
            x = r.ravel() / np.linalg.norm(r)
            assert np.abs(A @ x - ws[k] * x).max() < 1e-10


def test_double_seeds_follow_the_stable_diagonal_order():
    """The lowest pair-symmetric unit doubles are those of a stable sort of
    the diagonal, also where orbital energies are degenerate (ties)."""
    from helios_amd.lrccsd import _Layout, _Sub
    o, v = 3, 4
    for e in [np.array([-3.0, -2.0, -2.0, 1.0, 1.0, 2.5, 4.0]), np.array([-1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0])]:
        s = _Sub(o, v, e, 3, _Layout(3, 48))
        n1 = o * v
        m, n = np.triu_indices(n1)
        d = s._d2[m // v, n // v, m % v, n % v]
        for count in (1, 2, 5, len(d), len(d) + 3):
            want = []
            for q in np.argsort(d, kind="stable")[:count]:
                i, a, j, b = m[q] // v, m[q] % v, n[q] // v, n[q] % v
                want.append((n1 + ((i * o + j) * v + a) * v + b, n1 + ((j * o + i) * v + b) * v + a))
            assert s.double_seeds(count) == want


def test_average_excitation_level():
    rng = np.random.default_rng(3)
    o, v = 3, 4
    r1 = rng.standard_normal((o, v))
    r2 = rng.standard_normal((o, o, v, v))
    r2 = r2 + r2.transpose(1, 0, 3, 2)
    assert abs(lrccsd.average_excitation_level(r1, r2) - olr.average_excitation_level(r1, r2)) < 1e-12
    assert abs(lrccsd.average_excitation_level(r1, 0 * r2) - 1.0) < 1e-15
    assert abs(lrccsd.average_excitation_level(0 * r1, r2) - 2.0) < 1e-15


def _second_quantised_hbar(C, core, G, t1, t2):
    """e^-T H e^T for two electrons in the 2n spin orbitals (alpha, beta
    alternating), in the basis of all two-electron determinants, written
    directly from creation and annihilation operators on occupation tuples."""
    from itertools import combinations
    from oracle.triples import spin_amplitudes
    n = C.shape[0]
    ns = 2 * n
    dets = list(combinations(range(ns), 2))
    index = {d: k for k, d in enumerate(dets)}

    def annihilate(p, det):
        if p not in det:
            return None, 0
        pos = det.index(p)
        return det[:pos] + det[pos + 1:], (-1) ** pos

    def create(p, det):
        if p in det:
            return None, 0
        pos = sum(q < p for q in det)
        return tuple(sorted(det + (p,))), (-1) ** pos

    def apply(ops, det):
        """ops applied right to left: [(kind, p), ...] with kind '+' / '-'."""
        sign = 1
        for kind, p in reversed(ops):
            det, s = (create if kind == "+" else annihilate)(p, det)
            if det is None:
                return None, 0
            sign *= s
        return det, sign

    def matrix(terms):
        M = np.zeros((len(dets), len(dets)))
        for coef, ops in terms:
            if coef == 0.0:
                continue
            for k, d in enumerate(dets):
                out, s = apply(ops, d)
                if out is not None:
                    M[index[out], k] += coef * s
        return M

    h = C.T @ core @ C
    X = np.einsum("mp,mq->mpq", C, C)
    g = np.einsum("mpq,mn,nrs->pqrs", X, G, X, optimize=True)   # (pq|rs)
    sp, sg = np.arange(ns) // 2, np.arange(ns) % 2
    terms = []
    for p in range(ns):
        for q in range(ns):
            if sg[p] == sg[q]:
                terms.append((h[sp[p], sp[q]], [("+", p), ("-", q)]))
    for p in range(ns):
        for q in range(ns):
            for r in range(ns):
                for s in range(ns):
                    # 1/2 sum <pq|rs> a+p a+q a_s a_r, <pq|rs> = (pr|qs) with spins p~r, q~s
                    if sg[p] == sg[r] and sg[q] == sg[s]:
                        terms.append((0.5 * g[sp[p], sp[r], sp[q], sp[s]], [("+", p), ("+", q), ("-", s), ("-", r)]))
    H = matrix(terms)
    st1, st2 = spin_amplitudes(t1, t2)
    no = 2 * t1.shape[0]
    tt = []
    for i in range(no):
        for a in range(ns - no):
            tt.append((st1[i, a], [("+", no + a), ("-", i)]))
    for i in range(no):
        for j in range(no):
            for a in range(ns - no):
                for b in range(ns - no):
                    tt.append((0.25 * st2[i, j, a, b], [("+", no + a), ("+", no + b), ("-", j), ("-", i)]))
    T = matrix(tt)
    E = np.eye(len(dets))
    eT, emT = E + T + T @ T / 2, E - T + T @ T / 2      # T^3 = 0 for two electrons
    ref = index[(0, 1)]
    proj = {}
    for i in range(no):
        for a in range(ns - no):
            d, s = apply([("+", no + a), ("-", i)], (0, 1))
            proj[("s", i, a)] = (index[d], s)
    for i in range(no):
        for j in range(i + 1, no):
            for a in range(ns - no):
                for b in range(a + 1, ns - no):
                    d, s = apply([("+", no + a), ("+", no + b), ("-", j), ("-", i)], (0, 1))
                    if d is not None:
                        proj[("d", i, j, a, b)] = (index[d], s)
    return emT @ H @ eT, ref, proj


@pytest.mark.parametrize("mol", [M.polyene(4), M.random_fragment(6, seed=3)], ids=lambda m: m.name)
def test_reference_weight_two_electrons(mol):
    """For two electrons the EOM-CCSD right eigenvectors are eigenvectors of
    Hbar = e^-T H e^T in the full determinant space; their reference component
    is r0.  The reference's r0 formula (oracle) and its closed-shell form
    (helios_amd.lrccsd) evaluated on the singles and doubles of those
    eigenvectors reproduce it, and the eigenvalues minus E_CC are the roots of
    the closed-shell Jacobian.  Dropping the t1 dressing of Fia, or the
    exchange part of <ij||ab>, fails."""
    G, core, C, F, eref = _two_electrons(mol)
    r = ccsd.solve(C, F, G, 1, eref, eps=1e-12, diis_scale=True)
    t1, t2 = r["t1"], r["t2"]
    Hb, ref, proj = _second_quantised_hbar(C, core, G, t1, t2)
    lam, V = np.linalg.eig(Hb)
    lam, V = lam.real, V.real
    e0 = lam[np.argmax(np.abs(V[ref]))]            # CC ground state: the reference dominates
    jac, _ = olr.excitation_energies(C, F, G, 1, t1, t2)
    g = lrccsd.ovov_integrals(C, G, 1)
    v = C.shape[0] - 1
    checked = 0
    for k in np.argsort(lam):
        w = lam[k] - e0
        if w < 1e-6:
            continue
        x = V[:, k]
        r1 = np.array([[x[proj[("s", 0, 2 * a)][0]] * proj[("s", 0, 2 * a)][1] for a in range(v)]])
        if np.abs(r1).max() < 1e-3 or np.abs(jac - w).min() > 1e-7:
            continue      # triplets and states without singles weight
        r2 = np.zeros((1, 1, v, v))
        for a in range(v):
            for b in range(v):
                lo, hi = min(2 * a, 2 * b + 1), max(2 * a, 2 * b + 1)
                d, s = proj[("d", 0, 1, lo, hi)]
                r2[0, 0, a, b] = x[d] * s * (1 if 2 * a < 2 * b + 1 else -1)
        # the reference's normalisation applied to r0 (sign: arbitrary)
        norm = 2 * np.sum(r1 ** 2) + np.sum(r2 ** 2)
        want = abs(x[ref]) / np.sqrt(norm)
        got_oracle = olr.reference_weight(C, F, G, 1, t1, w, r1, r2)
        got = lrccsd.reference_weight(g, F[:1, 1:], t1, w, r1, r2)
        print(mol.name, "omega", w, "r0 Hbar", want, "oracle", got_oracle, "closed shell", got)
        assert abs(got_oracle - want) < 1e-8 and abs(got - want) < 1e-8
        checked += 1
    assert checked >= 3
