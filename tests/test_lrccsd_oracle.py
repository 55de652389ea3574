"""Singlet LR-CCSD oracle and host solver (CPU).

* Pin: for two electrons EOM-CCSD is exact, so every eigenvalue of the
  closed-shell singlet Jacobian (dimension ov + ov(ov+1)/2 = n(n+1)/2 - 1)
  equals an excitation energy of the two-electron singlet full CI, which is
  written independently in the MO product basis.
* The block Davidson of helios_amd.lrccsd, driven by the oracle's sigma
  (the same central difference the GPU computes), finds the three lowest
  real eigenvalues of the dense Jacobian, including the doubly excited
  ("dark") lowest state of the polyenes.
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
    """The calls helios_amd.lrccsd makes, answered by the oracle."""

    def __init__(self, items):
        self.items = items

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


def test_davidson_finds_the_lowest_roots():
    items = []
    for mol in [M.octatetraene(), M.polyene(6), M.random_fragment(10, seed=4)]:
        p = prepare(mol)
        r = ccsd.solve(p["C"], p["F"], p["G"], mol.nocc, p["eref"], eps=1e-11, diis_scale=True)
        items.append((p, mol.nocc, r["t1"], r["t2"]))
    om, res, its, conv = lrccsd.excitation_energies(_OracleCtx(items), [(o, p["n"] - o) for p, o, _, _ in items],
                                                    nroots=3)
    for (p, o, t1, t2), w, c in zip(items, om, conv):
        lam, _ = olr.excitation_energies(p["C"], p["F"], p["G"], o, t1, t2)
        print(p["n"], w, lam[:3])
        assert c
        assert np.abs(w - lam[:3]).max() < 2e-6
