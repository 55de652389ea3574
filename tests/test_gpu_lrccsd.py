"""Singlet LR-CCSD on the GPU (hcc_batch_lr_sigma + helios_amd.lrccsd)
against the oracle (oracle/lrccsd.py, pinned by the two-electron full CI):

* the device's sigma = J r at random pair-symmetric r equals the oracle's
  central difference of the same residual (1e-6 of max |sigma|), for every
  molecule of a batch, and the amplitudes are left as they were;
* the three lowest singlet excitation energies of a batch (polyenes with
  their doubly excited lowest state, honeycomb fragments) equal the lowest
  real eigenvalues of the oracle's dense Jacobian within 2e-6 eV;
* for two electrons they equal the singlet full-CI excitation energies;
* a molecule's roots in the batch equal its roots alone.
"""
import numpy as np
import pytest

from helios_amd import capi, lrccsd, molecules as M
from oracle import ccsd, lrccsd as olr, ppp, scf
from tests._molprep import prepare

pytestmark = pytest.mark.gpu


def _solved_ctx(preps):
    ctx = capi.Context(0, capi.HCC_FP64, 20)
    ctx.batch_load(preps)
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-10)
    _, _, st = ctx.batch_results()
    assert (st == 1).all()
    return ctx


def _pair_symmetric(rng, o, v):
    r2 = rng.standard_normal((o, o, v, v))
    return 0.5 * (r2 + r2.transpose(1, 0, 3, 2))


def test_sigma_matches_oracle():
    mols = [M.octatetraene(), M.random_fragment(12, seed=2), M.batch_member(11)]
    preps = [prepare(m) for m in mols]
    ctx = _solved_ctx(preps)
    rng = np.random.default_rng(5)
    r1s, r2s = [], []
    for m in mols:
        o, v = m.nocc, m.n - m.nocc
        r1s.append(rng.standard_normal((o, v)))
        r2s.append(_pair_symmetric(rng, o, v))
    before = ctx.batch_results()[0]
    s1, s2 = ctx.batch_lr_sigma(np.concatenate([x.ravel() for x in r1s]),
                                np.concatenate([x.ravel() for x in r2s]), 1e-5)
    a1 = a2 = 0
    for m, p, r1, r2 in zip(mols, preps, r1s, r2s):
        o, v = m.nocc, m.n - m.nocc
        ref = ccsd.solve(p["C"], p["F"], p["G"], o, p["eref"], eps=1e-11, diis_scale=True)
        q1, q2 = olr.sigma(p["C"], p["F"], p["G"], o, ref["t1"], ref["t2"], r1, r2, 1e-5)
        g1 = s1[a1:a1 + o * v].reshape(o, v)
        g2 = s2[a2:a2 + o * o * v * v].reshape(o, o, v, v)
        scale = max(np.abs(q1).max(), np.abs(q2).max())
        print(m.name, "max|dsigma|", np.abs(g1 - q1).max(), np.abs(g2 - q2).max(), "of", scale)
        assert np.abs(g1 - q1).max() < 1e-6 * scale and np.abs(g2 - q2).max() < 1e-6 * scale
        a1, a2 = a1 + o * v, a2 + o * o * v * v
    # amplitudes restored: a second sigma gives the same numbers, energies unchanged
    t1, t2 = ctx.batch_lr_sigma(np.concatenate([x.ravel() for x in r1s]), np.concatenate([x.ravel() for x in r2s]),
                                1e-5)
    assert np.array_equal(t1, s1) and np.array_equal(t2, s2)
    assert np.array_equal(ctx.batch_results()[0], before)
    ctx.close()


def test_lowest_roots_match_dense_jacobian():
    mols = [M.octatetraene(), M.polyene(6), M.random_fragment(10, seed=4), M.random_fragment(12, seed=2)]
    preps = [prepare(m) for m in mols]
    ctx = _solved_ctx(preps)
    om, res, its, conv = lrccsd.excitation_energies(ctx, [(m.nocc, m.n - m.nocc) for m in mols], nroots=3)
    ctx.close()
    for m, p, w, c, it in zip(mols, preps, om, conv, its):
        r = ccsd.solve(p["C"], p["F"], p["G"], m.nocc, p["eref"], eps=1e-11, diis_scale=True)
        lam, _ = olr.excitation_energies(p["C"], p["F"], p["G"], m.nocc, r["t1"], r["t2"])
        print(m.name, "gpu", w, "oracle", lam[:3], "iterations", it, "diff", np.abs(w - lam[:3]).max())
        assert c
        assert np.abs(w - lam[:3]).max() < 2e-6


def test_two_electrons_equal_fci():
    mols = [M.ethylene(), M.polyene(4), M.random_fragment(6, seed=3)]
    preps, fcis = [], []
    for m in mols:
        G, core, hk = ppp.build(m.coords, m.bonds, alpha=m.alpha, U=m.U, kind="ohno-klopman")
        r = scf.rhf(core, G, hk, 1)
        F, eref = scf.mo_fock(r["C"], core, G, 1)
        preps.append(dict(C=r["C"], F=F, G=G, nocc=1, eref=eref, n=m.n))
        fcis.append(olr.singlets_two_electrons(r["C"], core, G))
    ctx = _solved_ctx(preps)
    om, res, its, conv = lrccsd.excitation_energies(ctx, [(1, m.n - 1) for m in mols], nroots=2)
    ctx.close()
    for m, w, f, c in zip(mols, om, fcis, conv):
        print(m.name, "gpu", w, "fci", f[:2])
        assert c
        assert np.abs(w - f[:2]).max() < 2e-6


def test_batch_equals_alone():
    mols = [M.polyene(6), M.random_fragment(10, seed=4)]
    preps = [prepare(m) for m in mols]
    ctx = _solved_ctx(preps)
    om, _, _, _ = lrccsd.excitation_energies(ctx, [(m.nocc, m.n - m.nocc) for m in mols], nroots=2)
    ctx.close()
    for k, (m, p) in enumerate(zip(mols, preps)):
        one = _solved_ctx([p])
        w1, _, _, _ = lrccsd.excitation_energies(one, [(m.nocc, m.n - m.nocc)], nroots=2)
        one.close()
        assert np.abs(w1[0] - om[k]).max() < 1e-6
