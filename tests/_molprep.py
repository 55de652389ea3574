"""Shared test helper: PPP model + oracle SCF for a molecule."""
import numpy as np

from oracle import ppp, scf


def prepare(mol, kind="ohno-klopman"):
    G, core, hk = ppp.build(mol.coords, mol.bonds, alpha=mol.alpha, U=mol.U, kind=kind)
    r = scf.rhf(core, G, hk, mol.nocc)
    F, eref = scf.mo_fock(r["C"], core, G, mol.nocc)
    return dict(G=G, core=core, huckel=hk, C=r["C"], E=r["E"], F=F, eref=eref,
                nocc=mol.nocc, n=mol.n, scf=r)


def random_amplitudes(o, v, seed, s1=0.1, s2=0.1):
    rng = np.random.default_rng(seed)
    t1 = s1 * rng.standard_normal((o, v))
    t2 = s2 * rng.standard_normal((o, o, v, v))
    return t1, 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
