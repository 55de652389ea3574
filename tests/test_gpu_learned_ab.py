"""What the trained network gains over the MP2 guess (A/B on held-out molecules).

The committed parameters (helios_amd/data/guess_mlp.f32) were trained on
20000 molecules drawn with seeds 1000000 + k; the test molecules are the first
256 members of the benchmark batch (seeds 3 + k), so no test molecule was
seen in training.  Both guesses are solved with the plain FP64 solver to the
same residual threshold; every solve must converge to within 1e-6 eV of a
tight (1e-10) FP64 reference, and the learned guess must need fewer
iterations in total.
"""
import numpy as np
import pytest

from helios_amd import capi, learned_guess, molecules as M, ppp

pytestmark = pytest.mark.gpu


def batch(count):
    out = []
    for k in range(count):
        p = ppp.prepare(M.batch_member(k))
        out.append(dict(C=p["C"], F=p["F"], G=p["G"], nocc=p["nocc"], eref=p["eref"]))
    return out


def test_learned_guess_cuts_iterations():
    mols = batch(256)
    params = learned_guess.load_params()
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    ctx.set_guess_network(params)
    ctx.batch_load(mols)
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-10)
    eref, _, sref = ctx.batch_results()
    assert (sref == 1).all()
    its = {}
    for kind in (capi.GUESS_MP2, capi.GUESS_LEARNED):
        ctx.batch_solve(kind, 500, 2e-7)
        e, it, st = ctx.batch_results()
        assert (st == 1).all()
        assert np.abs(e[:, 0] - eref[:, 0]).max() < 1e-6
        its[kind] = it.astype(np.float64)
    saved = its[capi.GUESS_MP2].sum() - its[capi.GUESS_LEARNED].sum()
    print(f"mean iterations: MP2 {its[capi.GUESS_MP2].mean():.2f}, learned {its[capi.GUESS_LEARNED].mean():.2f}; "
          f"fewer for {(its[capi.GUESS_LEARNED] < its[capi.GUESS_MP2]).sum()} of {len(mols)} molecules")
    assert saved > 0.01 * its[capi.GUESS_MP2].sum()
