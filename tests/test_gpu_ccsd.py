"""Parity of the HIP r-ccsd solver with the oracle, through the C ABI.

Tolerances: energies within 1e-8 eV and T2 within 1e-7 max-abs of the
oracle's converged FP64 solution on the same inputs (C, F_mo, gamma, e_ref);
residuals at random amplitudes within 1e-11 relative (FP64 summation-order
differences only).
"""
import numpy as np
import pytest

from helios_amd import capi, molecules as M
from oracle import ccsd
from tests._molprep import prepare, random_amplitudes

pytestmark = pytest.mark.gpu


def context(p, diis_steps=20):
    ctx = capi.Context(0, capi.HCC_FP64, diis_steps)
    ctx.set_molecule(p["C"], p["F"], p["G"], p["nocc"], p["eref"])
    return ctx


@pytest.mark.parametrize("name", ["ethylene", "octatetraene", "polyene10", "frag"])
def test_residual_parity_random_amplitudes(name):
    mol = {"ethylene": M.ethylene(), "octatetraene": M.octatetraene(), "polyene10": M.polyene(10),
           "frag": M.batch_member(4)}[name]
    if mol.n > 20:
        mol = M.random_fragment(14, seed=1)
    p = prepare(mol)
    o, v = p["nocc"], p["n"] - p["nocc"]
    ctx = context(p)
    B = ccsd.Blocks(ccsd.rint(p["C"], p["G"]), o)
    for seed in (0, 1):
        t1, t2 = random_amplitudes(o, v, seed)
        ctx.set_amplitudes(t1, t2)
        g1, g2 = ctx.amplitudes()
        assert np.array_equal(g1, t1) and np.array_equal(g2, t2)
        d1, d2 = ctx.residual()
        r1 = ccsd.singles(p["F"], B, t1, t2)
        r2 = ccsd.doubles(p["F"], B, t1, t2)
        s1 = max(1.0, np.abs(r1).max())
        s2 = max(1.0, np.abs(r2).max())
        print(name, seed, np.abs(d1 - r1).max() / s1, np.abs(d2 - r2).max() / s2)
        assert np.abs(d1 - r1).max() < 1e-11 * s1
        assert np.abs(d2 - r2).max() < 1e-11 * s2
    ctx.close()


@pytest.mark.parametrize("name", ["ethylene", "octatetraene", "frag10", "frag16"])
def test_solve_parity(name):
    mol = {"ethylene": M.ethylene(), "octatetraene": M.octatetraene(),
           "frag10": M.random_fragment(10, seed=3), "frag16": M.random_fragment(16, seed=5)}[name]
    p = prepare(mol)
    o = p["nocc"]
    ref = ccsd.solve(p["C"], p["F"], p["G"], o, p["eref"], eps=1e-11, diis_scale=True)
    assert ref["status"] == 1
    ctx = context(p)
    e, it, st = ctx.solve(capi.GUESS_REFERENCE, 2000, 1e-11)
    t1, t2 = ctx.amplitudes()
    ctx.close()
    print(name, "E", e[0], ref["energy"][0], "iters", it, ref["iterations"])
    assert st == 1
    assert abs(e[0] - ref["energy"][0]) < 1e-8
    assert abs(e[1] - p["eref"]) < 1e-12
    assert np.abs(t2 - ref["t2"]).max() < 1e-7
    assert np.abs(t1 - ref["t1"]).max() < 1e-7


def test_iterate_mirrors_iterationCC():
    """hcc_iterate + hcc_energy step for step against the oracle's iterationCC."""
    p = prepare(M.octatetraene())
    o = p["nocc"]
    cc = ccsd.CCSD(p["C"], p["F"], p["G"], o, p["eref"], diis_scale=True)
    ctx = context(p)
    ctx.guess(capi.GUESS_REFERENCE)
    t1, t2 = ctx.amplitudes()
    assert np.abs(t2 - cc.t2).max() < 1e-12
    for it in range(1, 9):
        acc = ctx.iterate(it, 1e-12)
        ref_acc = cc.iterate(it, 1e-12)
        assert abs(acc[0] - ref_acc[0]) < 1e-9 * max(1, ref_acc[0])
        assert abs(acc[1] - ref_acc[1]) < 1e-9 * max(1, ref_acc[1])
        e = ctx.energy()
        assert abs(e[0] - cc.energy()[0]) < 1e-9
    ctx.close()


def test_batch_matches_single_solves():
    mols = [M.ethylene(1.40), M.random_fragment(10, seed=11), M.octatetraene(), M.random_fragment(12, seed=2)]
    ps = [prepare(m) for m in mols]
    ctx = capi.Context(0, capi.HCC_FP64, 20)
    ctx.batch_load(ps)
    ctx.batch_solve(capi.GUESS_REFERENCE, 500, 1e-11)
    e, it, st = ctx.batch_results()
    ctx.close()
    for k, p in enumerate(ps):
        ref = ccsd.solve(p["C"], p["F"], p["G"], p["nocc"], p["eref"], eps=1e-11, diis_scale=True)
        assert st[k] == 1
        assert abs(e[k, 0] - ref["energy"][0]) < 1e-8


@pytest.mark.parametrize("precision", [capi.HCC_FP64, capi.HCC_MIXED])
def test_batch_members_equal_alone_bitwise(precision):
    """Molecules that converge at different iterations: once some have
    finished, the GEMM tables are compacted to the running ones; every
    member's energies, iteration count and status equal its solve alone
    bitwise (a molecule dropped from a table too early, or a tile
    renumbered wrongly, would change them)."""
    ps = [prepare(m) for m in [M.ethylene(1.40), M.random_fragment(10, seed=11), M.octatetraene(),
                               M.random_fragment(12, seed=2), M.batch_member(5), M.batch_member(17)]]
    ctx = capi.Context(0, precision, 8)
    ctx.batch_load(ps)
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-9)
    e, it, st = ctx.batch_results()
    ctx.close()
    assert (st == 1).all() and len(set(it.tolist())) > 2
    for k, p in enumerate(ps):
        one = capi.Context(0, precision, 8)
        one.batch_load([p])
        one.batch_solve(capi.GUESS_MP2, 500, 1e-9)
        e1, it1, st1 = one.batch_results()
        one.close()
        print(k, it[k], it1[0], e[k, 0], e1[0, 0])
        assert np.array_equal(e1[0], e[k]) and it1[0] == it[k] and st1[0] == st[k]


@pytest.mark.parametrize("precision", [capi.HCC_FP64, capi.HCC_MIXED])
def test_graph_replay_equals_eager_launches(precision):
    """Iteration blocks replayed from captured HIP graphs give bitwise the
    results of launching every kernel from the host, solve after solve (each
    solve replays its graph many times, with host synchronisations between)."""
    ps = [prepare(M.batch_member(k)) for k in range(12)]
    out = []
    for graphs in (False, True):
        ctx = capi.Context(0, precision, 8)
        ctx.set_graphs(graphs)
        ctx.batch_load(ps)
        runs = []
        for _ in range(2):
            ctx.batch_solve(capi.GUESS_MP2, 500, 1e-9)
            runs.append(ctx.batch_results())
        ctx.close()
        for r in runs[1:]:
            assert all(np.array_equal(a, b) for a, b in zip(runs[0], r))
        out.append(runs[0])
    e0, it0, st0 = out[0]
    e1, it1, st1 = out[1]
    assert (st0 == 1).all()
    assert np.array_equal(it0, it1) and np.array_equal(st0, st1)
    assert np.array_equal(e0, e1)
