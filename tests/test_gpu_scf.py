"""Batched RHF on the GPU (hcc_batch_load_scf) against the oracle's
restatement of the reference SCF (oracle/scf.py: Hueckel guess, damped-density
steps of 0.05, src/methods/mod_scf.f90:106-136, 351-386, 446-494):

* per molecule: the same number of iterations (+-1: the stopping test reads
  a last-digit quantity), reference energy within 1e-9 eV, MO coefficients
  (columns up to sign) and MO Fock matrix (elements up to the signs of their
  two orbitals) within 1e-7;
* the CCSD energies solved from the device SCF equal those from the host SCF
  path within 1e-8 eV;
* a molecule beyond the device SCF's size limit is rejected with an error.
"""
import numpy as np
import pytest

from helios_amd import capi, molecules as M, ppp
from oracle import scf

pytestmark = pytest.mark.gpu


def scf_inputs(mol):
    G, core, hk = ppp.model(mol.coords, mol.bonds, mol.alpha, mol.U, 1)
    return dict(core=core, G=G, huckel=hk, nocc=mol.nocc, n=mol.n)


MOLS = [M.octatetraene(), M.polyene(20), M.ethylene(1.40)] + [M.batch_member(k) for k in range(12)] + \
       [M.random_fragment(48, seed=31)]


def test_device_scf_matches_oracle():
    ins = [scf_inputs(m) for m in MOLS]
    ctx = capi.Context(0, capi.HCC_FP64, 20)
    er, it, st = ctx.batch_load_scf(ins, accuracy=1e-10, maxiter=400)
    worst = [0.0, 0.0, 0.0]
    for k, (m, p) in enumerate(zip(MOLS, ins)):
        ro = scf.rhf(p["core"], p["G"], p["huckel"], m.nocc, accuracy=1e-10, maxiter=400)
        Fo, eo = scf.mo_fock(ro["C"], p["core"], p["G"], m.nocc)
        C, F, e = ctx.batch_orbitals(k, m.n)
        # columns up to sign: in a symmetric molecule the largest |C| of an
        # orbital sits on symmetry-equivalent sites, and the gauge convention
        # (largest entry positive) depends on rounding there
        dc = float(np.minimum(np.abs(C - ro["C"]).max(0), np.abs(C + ro["C"]).max(0)).max())
        df = np.abs(np.abs(F) - np.abs(Fo)).max()
        worst = [max(worst[0], abs(e - eo)), max(worst[1], dc), max(worst[2], df)]
        print(m.name, m.n, "iterations", it[k], ro["iterations"], "dE", e - eo, "dC", dc, "dF", df)
        assert st[k] == 1 and ro["converged"]
        assert abs(int(it[k]) - ro["iterations"]) <= 1
        assert abs(er[k] - e) == 0.0
        assert abs(e - eo) < 1e-9
        assert dc < 1e-7 and df < 1e-7
    ctx.close()
    print("worst dE, dC, dF", worst)


def test_ccsd_from_device_scf_matches_host_scf():
    mols = MOLS[:1] + [M.batch_member(k) for k in range(40)]
    dev = capi.Context(0, capi.HCC_FP64, 20)
    dev.batch_load_scf([scf_inputs(m) for m in mols], accuracy=1e-11, maxiter=400)
    dev.batch_solve(capi.GUESS_MP2, 500, 1e-10)
    e_dev, _, st_dev = dev.batch_results()
    dev.close()
    host = capi.Context(0, capi.HCC_FP64, 20)
    host.batch_load([ppp.prepare(m) for m in mols])
    host.batch_solve(capi.GUESS_MP2, 500, 1e-10)
    e_host, _, st_host = host.batch_results()
    host.close()
    d = np.abs(e_dev[:, 0] - e_host[:, 0])
    print("CCSD from device SCF vs host SCF: max |dE|", d.max())
    assert (st_dev == 1).all() and (st_host == 1).all()
    assert d.max() < 1e-8


def test_device_scf_rejects_oversize():
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    with pytest.raises(capi.HccError, match="up to 64 sites"):
        ctx.batch_load_scf([scf_inputs(M.random_fragment(70, seed=1))])
    ctx.close()
