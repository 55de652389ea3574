"""The C restatement of the reference CPU path (oracle/refloops.c, the CPU
baseline of bench.py) against the numpy oracle: the R build, both
projections at random amplitudes (term by term the same sums, so agreement to
rounding), and a whole solve (energy, iterations, amplitudes), with one and
several OpenMP threads."""
import os

import numpy as np
import pytest

from helios_amd import molecules as M
from oracle import ccsd, refloops
from tests._molprep import prepare, random_amplitudes

pytestmark = pytest.mark.skipif(not os.path.exists(refloops.LIB_PATH), reason="make -C oracle not run")


def test_R_build_matches_oracle():
    p = prepare(M.octatetraene())
    assert np.abs(refloops.build_R(p["C"], p["G"]) - ccsd.rint(p["C"], p["G"])).max() < 1e-12


@pytest.mark.parametrize("name", ["octatetraene", "frag10", "frag12"])
def test_projections_match_oracle(name):
    mol = {"octatetraene": M.octatetraene(), "frag10": M.random_fragment(10, seed=4),
           "frag12": M.random_fragment(12, seed=9)}[name]
    p = prepare(mol)
    o, v = p["nocc"], p["n"] - p["nocc"]
    B = ccsd.Blocks(ccsd.rint(p["C"], p["G"]), o)
    t1, t2 = random_amplitudes(o, v, 5)
    d1, d2 = refloops.residuals(p["C"], p["G"], p["F"], o, t1, t2)
    r1 = ccsd.singles(p["F"], B, t1, t2)
    r2 = ccsd.doubles(p["F"], B, t1, t2)
    assert np.abs(d1 - r1).max() < 1e-12 * max(1.0, np.abs(r1).max())
    assert np.abs(d2 - r2).max() < 1e-12 * max(1.0, np.abs(r2).max())


@pytest.mark.parametrize("threads", [1, 4])
def test_solve_matches_oracle(threads):
    p = prepare(M.octatetraene())
    ref = ccsd.solve(p["C"], p["F"], p["G"], 4, p["eref"], eps=1e-10, diis_scale=True)
    r = refloops.solve(p["C"], p["F"], p["G"], 4, p["eref"], eps=1e-10, threads=threads)
    assert r["status"] == 1
    assert abs(r["energy"][0] - ref["energy"][0]) < 1e-10
    assert abs(r["iterations"] - ref["iterations"]) <= 1
    assert np.abs(r["t2"] - ref["t2"]).max() < 1e-8
