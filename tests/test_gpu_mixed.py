"""Mixed-precision (defect-correction) solver against the FP64 solver.

* exact-correction mode (correction-form GEMMs kept in FP64, corrections
  stored in FP32, tight inner tolerance): every correction-form step is the
  exact change of its FP64 counterpart up to the FP32 rounding of the
  corrections, so two inner solves reach the FP64 fixed point and the third
  FP64 residual evaluation confirms it -- a wrong correction step shows up as
  extra outer evaluations or a different energy;
* BF16 mode: energies within 1e-6 eV of the FP64 energies of the same inputs,
  with the inner DIIS history restarted or kept across outer points;
* kept history in exact mode: the stored corrections are shifted at every
  commit, so the kept vectors stay consistent -- the solve reaches the same
  fixed point and needs no more inner iterations than with a restarted
  history (a wrong shift extrapolates from inconsistent vectors).
"""
import numpy as np
import pytest

from helios_amd import capi, molecules as M, ppp

pytestmark = pytest.mark.gpu


def batch(ks):
    return [ppp.prepare(M.batch_member(k)) for k in ks] + [ppp.prepare(M.octatetraene())]


def fp64_energies(mols, eps=1e-10):
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    ctx.batch_load(mols)
    ctx.batch_solve(capi.GUESS_MP2, 500, eps)
    e, it, st = ctx.batch_results()
    ctx.close()
    assert (st == 1).all()
    return e[:, 0], it


def test_exact_correction_reaches_fixed_point():
    mols = batch([3, 7, 12])
    ref, _ = fp64_energies(mols)
    ctx = capi.Context(0, capi.HCC_MIXED, 8)
    ctx.set_mixed(bf16=False, kappa=1e-5, max_inner=300, max_outer=10)
    ctx.batch_load(mols)
    ctx.batch_solve(capi.GUESS_MP2, 2000, 1e-8)
    e, it, st = ctx.batch_results()
    outer, inner = ctx.batch_statistics()
    ctx.close()
    print("outer", outer, "inner", inner, "dE", e[:, 0] - ref)
    assert (st == 1).all()
    assert (outer <= 3).all()
    assert np.abs(e[:, 0] - ref).max() < 1e-8


def test_kept_history_exact_mode():
    mols = batch([3, 7, 12, 20])
    ref, _ = fp64_energies(mols)
    inner_total = {}
    for keep in (False, True):
        ctx = capi.Context(0, capi.HCC_MIXED, 8)
        ctx.set_mixed(bf16=False, kappa=0.1, max_inner=40, max_outer=30)
        ctx.set_mixed_history(keep)
        ctx.batch_load(mols)
        ctx.batch_solve(capi.GUESS_MP2, 2000, 1e-8)
        e, it, st = ctx.batch_results()
        outer, inner = ctx.batch_statistics()
        ctx.close()
        print("keep", keep, "outer", outer, "inner", inner, "dE", e[:, 0] - ref)
        assert (st == 1).all()
        # residuals converged to 1e-8 leave energies a few 1e-9 from the reference
        assert np.abs(e[:, 0] - ref).max() < 3e-8
        inner_total[keep] = int(inner.sum())
    assert inner_total[True] <= inner_total[False]


@pytest.mark.parametrize("keep", [False, True])
def test_bf16_within_1e6_ev(keep):
    mols = batch([3, 4, 5, 7, 12, 20])
    ref, it64 = fp64_energies(mols)
    ctx = capi.Context(0, capi.HCC_MIXED, 8)
    ctx.set_mixed(bf16=True, kappa=0.1, max_inner=40, max_outer=20)
    ctx.set_mixed_history(keep)
    ctx.batch_load(mols)
    ctx.batch_solve(capi.GUESS_MP2, 2000, 2e-7)
    e, it, st = ctx.batch_results()
    outer, inner = ctx.batch_statistics()
    ctx.close()
    print("fp64 its", it64, "outer", outer, "inner", inner, "dE", e[:, 0] - ref)
    assert (st == 1).all()
    assert np.abs(e[:, 0] - ref).max() < 1e-6


@pytest.mark.parametrize("tile,ksplit", [(32, 1), (64, 2), (256, 1)])
def test_tuned_correction_tiles(tile, ksplit):
    """Tuning entries for the correction forms (keys 100 + op) change the
    tiles and K splits of the BF16 correction GEMMs; the solve still reaches
    the FP64 energies within 1e-6 eV.  An FP64-only tile kind is refused for
    a correction key, and correction timing needs a mixed context."""
    mols = batch([3, 5, 12, 20])
    ref, _ = fp64_energies(mols)
    sizes = sorted({(m["nocc"], m["C"].shape[0] - m["nocc"]) for m in mols})
    entries = [[100 + op, o, v, tile, ksplit] for op in range(25) for o, v in sizes]
    ctx = capi.Context(0, capi.HCC_MIXED, 8)
    ctx.set_gemm_tuning(entries)
    ctx.batch_load(mols)
    assert ctx.tune_gemm(100 + 3, tile, ksplit, 2) > 0.0
    ctx.batch_solve(capi.GUESS_MP2, 2000, 2e-7)
    e, it, st = ctx.batch_results()
    with pytest.raises(capi.HccError):
        ctx.set_gemm_tuning([[100, 5, 5, 128, 1]])
    ctx.close()
    print(tile, ksplit, "dE", e[:, 0] - ref)
    assert (st == 1).all()
    assert np.abs(e[:, 0] - ref).max() < 1e-6
    f64 = capi.Context(0, capi.HCC_FP64, 8)
    f64.batch_load(mols)
    with pytest.raises(capi.HccError):
        f64.tune_gemm(100 + 3, 64, 1, 1)
    f64.close()
