"""hcc_solve_noalloc (the solve whose device buffers all come from the
library's static scratch region) against hcc_solve through a context, to the
parity bar of the FP64 solver: two calls in a row, the second on a smaller
molecule than the first, each compared with the context solve of the same
molecule; then a molecule too large for the region is rejected with a message
and a later call still works."""
import numpy as np
import pytest

from helios_amd import capi, molecules as M
from tests._molprep import prepare

pytestmark = pytest.mark.gpu


def context_solve(p, guess, eps, diis):
    ctx = capi.Context(0, capi.HCC_FP64, diis)
    ctx.set_molecule(p["C"], p["F"], p["G"], p["nocc"], p["eref"])
    e, it, st = ctx.solve(guess, 2000, eps)
    t1, t2 = ctx.amplitudes()
    ctx.close()
    return e, it, st, t1, t2


@pytest.mark.parametrize("guess", [capi.GUESS_REFERENCE, capi.GUESS_MP2])
def test_noalloc_matches_context_solve(guess):
    mols = [prepare(M.random_fragment(48, seed=21)), prepare(M.random_fragment(16, seed=22))]
    for p in mols:     # larger first, then smaller: the region keeps the first call's contents
        e, it, st, t1, t2 = capi.solve_noalloc(p["C"], p["F"], p["G"], p["nocc"], p["eref"], guess=guess,
                                               eps=1e-10, diis_steps=20)
        er, itr, str_, t1r, t2r = context_solve(p, guess, 1e-10, 20)
        print(p["n"], "noalloc", e[0], it, st, "context", er[0], itr, str_)
        assert st == 1 and str_ == 1 and it == itr
        assert np.abs(e - er).max() < 1e-10
        assert np.abs(t2 - t2r).max() < 1e-10 and np.abs(t1 - t1r).max() < 1e-10


def test_noalloc_rejects_molecule_beyond_region():
    p = prepare(M.random_fragment(64, seed=23))
    with pytest.raises(capi.HccError, match="no-allocation scratch region"):
        capi.solve_noalloc(p["C"], p["F"], p["G"], p["nocc"], p["eref"], eps=1e-8)
    q = prepare(M.octatetraene())
    e, it, st, _, _ = capi.solve_noalloc(q["C"], q["F"], q["G"], q["nocc"], q["eref"], eps=1e-10)
    er, itr, _, _, _ = context_solve(q, capi.GUESS_MP2, 1e-10, 20)
    assert st == 1 and it == itr and abs(e[0] - er[0]) < 1e-10
