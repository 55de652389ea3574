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

    def batch_orbitals(self, k, n):
        p = self.items[k][0]
        return p["C"], p["F"], p["eref"]

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
    om, res, its, conv, ael = lrccsd.excitation_energies(
        _OracleCtx(items), [(o, p["n"] - o) for p, o, _, _ in items], nroots=3, gammas=gammas)
    for (p, o, t1, t2), w, c, lv in zip(items, om, conv, ael):
        lam, _ = olr.excitation_energies(p["C"], p["F"], p["G"], o, t1, t2)
        print(p["n"], w, lam[:3], "AEL", lv)
        assert c
        assert np.abs(w - lam[:3]).max() < 2e-6
        assert np.all((lv >= 1.0) & (lv <= 2.0))


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


def test_average_excitation_level():
    rng = np.random.default_rng(3)
    o, v = 3, 4
    r1 = rng.standard_normal((o, v))
    r2 = rng.standard_normal((o, o, v, v))
    r2 = r2 + r2.transpose(1, 0, 3, 2)
    assert abs(lrccsd.average_excitation_level(r1, r2) - olr.average_excitation_level(r1, r2)) < 1e-12
    assert abs(lrccsd.average_excitation_level(r1, 0 * r2) - 1.0) < 1e-15
    assert abs(lrccsd.average_excitation_level(0 * r1, r2) - 2.0) < 1e-15
