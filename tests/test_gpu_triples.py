"""(T) correction on the GPU (hcc_batch_triples) against the oracle
(oracle/triples.py, itself pinned to the restatement of the reference's
spin-orbital loop, src/methods/coupledCluster/mod_cc.f90:2028-2111):

* kernel parity at given amplitudes: one molecule with random amplitudes set
  through hcc_set_amplitudes; E[T] and E[ST] equal the oracle's within 1e-11
  of |E[T]|, at o = v = 10 and 20 and at both size limits (32 occupied with
  8 virtual, 8 occupied with 32 virtual); the amplitudes are left untouched
  and a second call gives the same numbers;
* converged amplitudes, ragged batch: E[T] + E[ST] of every molecule within
  1e-8 eV of the oracle's on the oracle's own CCSD amplitudes, and the
  reference's energy(1..5) of 'spin-r-ccsd(t)' within 1e-8 eV of the
  restated loop on the small ones;
* a molecule's result in the batch equals its result solved alone;
* a molecule with 33 occupied orbitals is rejected.
"""
import numpy as np
import pytest

from helios_amd import capi, molecules as M
from oracle import ccsd, triples as T
from tests._molprep import prepare, random_amplitudes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nocc", [(20, 10), (40, 20), (40, 32), (40, 8)])
def test_kernel_matches_oracle_at_given_amplitudes(n, nocc):
    p = prepare(M.random_fragment(n, seed=11))
    o, v = nocc, n - nocc
    t1, t2 = random_amplitudes(o, v, 100 + n + nocc, 0.05, 0.05)
    ctx = capi.Context(0, capi.HCC_FP64, 20)
    ctx.set_molecule(p["C"], p["F"], p["G"], o, p["eref"])
    ctx.set_amplitudes(t1, t2)
    e, et = ctx.batch_triples()
    a1, a2 = ctx.amplitudes()
    e2, et2 = ctx.batch_triples()
    ctx.close()
    ref = T.closed_shell_sliced(p["C"], p["F"], p["G"], o, t1, t2)
    print(n, o, v, "E[T]", et[0], "oracle", ref, "diff", et[0] - ref)
    assert abs(et[0, 0] - ref[0]) < 1e-11 * abs(ref[0])
    assert abs(et[0, 1] - ref[1]) < 1e-11 * abs(ref[0])
    assert np.array_equal(a1, t1) and np.array_equal(a2, t2)
    assert np.array_equal(et, et2) and np.array_equal(e, e2)
    assert e[0, 4] == -et[0, 0]


SMALL = [M.octatetraene(), M.polyene(6)]
BATCH = SMALL + [M.batch_member(k) for k in (8, 11, 18, 3)] + [M.random_fragment(24, seed=7)]


def test_converged_batch_matches_oracle():
    preps = [prepare(m) for m in BATCH]
    ctx = capi.Context(0, capi.HCC_FP64, 20)
    ctx.batch_load(preps)
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-10)
    eb, _, st = ctx.batch_results()
    e, et = ctx.batch_triples()
    eb_after, _, _ = ctx.batch_results()
    ctx.close()
    assert (st == 1).all()
    assert np.array_equal(eb, eb_after)
    for k, (m, p) in enumerate(zip(BATCH, preps)):
        r = ccsd.solve(p["C"], p["F"], p["G"], m.nocc, p["eref"], eps=1e-11, diis_scale=True)
        assert r["status"] == 1
        eT, eST = T.closed_shell_sliced(p["C"], p["F"], p["G"], m.nocc, r["t1"], r["t2"])
        print(m.name, m.n, "E(T) gpu", et[k].sum(), "oracle", eT + eST, "diff", et[k].sum() - (eT + eST))
        assert abs(et[k].sum() - (eT + eST)) < 1e-8
        assert abs(e[k, 0] - (eb[k, 0] - et[k].sum())) < 1e-12 * abs(eb[k, 0])
        assert e[k, 1] == eb[k, 1] and e[k, 2] == eb[k, 2] and e[k, 3] == eb[k, 3]
        if k < len(SMALL):
            c1, c2, c3 = T.triples(p["C"], p["F"], p["G"], m.nocc, r["t1"], r["t2"])
            ref = np.array([r["energy"][0] + c1 + c2 + c3, r["energy"][1], r["energy"][2], r["energy"][3], c1])
            print("  energy(1..5)", e[k], "restated loop", ref)
            assert np.abs(e[k] - ref).max() < 1e-8


def test_batch_member_equals_alone():
    mols = [M.batch_member(k) for k in (3, 8, 11)]
    preps = [prepare(m) for m in mols]
    ctx = capi.Context(0, capi.HCC_FP64, 20)
    ctx.batch_load(preps)
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-11)
    _, et = ctx.batch_triples()
    ctx.close()
    for k, p in enumerate(preps):
        one = capi.Context(0, capi.HCC_FP64, 20)
        one.batch_load([p])
        one.batch_solve(capi.GUESS_MP2, 500, 1e-11)
        _, et1 = one.batch_triples()
        one.close()
        assert np.abs(et1[0] - et[k]).max() < 1e-10 * abs(et[k, 0])


def test_rejects_more_than_32_occupied():
    p = prepare(M.random_fragment(40, seed=3))
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    ctx.set_molecule(p["C"], p["F"], p["G"], 33, p["eref"])
    ctx.set_amplitudes(*random_amplitudes(33, 7, 1))
    with pytest.raises(capi.HccError, match="up to 32 occupied"):
        ctx.batch_triples()
    ctx.close()


def test_batch_across_both_kernels():
    """Molecules with v <= 24 take the matrix-core kernel (W in LDS), larger v
    the global-W kernel; in one batch each skips the other's work items."""
    mols = [M.octatetraene(), M.random_fragment(52, seed=5), M.batch_member(8)]
    preps = [prepare(m) for m in mols]
    assert preps[1]["n"] - preps[1]["nocc"] == 26
    ctx = capi.Context(0, capi.HCC_FP64, 20)
    ctx.batch_load(preps)
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-10)
    _, et = ctx.batch_triples()
    ctx.close()
    for k, p in enumerate(preps):
        one = capi.Context(0, capi.HCC_FP64, 20)
        one.batch_load([p])
        one.batch_solve(capi.GUESS_MP2, 500, 1e-10)
        _, et1 = one.batch_triples()
        one.close()
        print(mols[k].name, et[k], et1[0])
        assert np.abs(et1[0] - et[k]).max() < 1e-10 * abs(et[k, 0])
