"""Finite-field polarizability of the reference, restated (test
infrastructure only): the field perturbation of the core (setCore,
src/modules/mod_property.f90:467-486, with the centred coordinates of
src/input/readMoleculeInformation.f90:381-397 and perturbate,
src/modules/mod_helios.f90:577-601), the oracle SCF and CCSD at every point,
and the reference's Lagrange derivative (LagrangeDerivative,
src/modules/mod_math.f90:505-560, built from elementary symmetric products of
the other nodes as the reference does).

Pinning (tests/test_polarizability.py): the derivative is exact on
polynomials of degree below the number of points and agrees with an
independent Vandermonde solve; for two electrons the CCSD energies at every
field point equal the determinant FCI's (oracle.fci), so the CCSD
polarizability equals the FCI one.
"""
import math
from itertools import combinations

import numpy as np

from oracle import ccsd, ppp, scf


def set_core(core, huckel, coords, point, step):
    """core + P and huckel + P, P = diag(F . r_mu), F = point * step."""
    r = np.asarray(coords, dtype=np.float64)
    r = r - r.mean(0)
    P = np.diag(r @ (np.asarray(point, dtype=np.float64) * step))
    return core + P, huckel + P


def lagrange_derivative(nump, power, grid, ff):
    """LagrangeDerivative(nump, power, enGrid, ff): power-th derivative at 0 of
    the Lagrange polynomial through enGrid at nodes -s..s (spacing ff).  The
    coefficient of x^power in prod_{k != j} (x - k) is, up to sign, the
    elementary symmetric sum of degree nump-1-power of the other nodes."""
    s = (nump - 1) // 2
    nodes = list(range(-s, s + 1))
    total = 0.0
    for idx, j in enumerate(nodes):
        others = [k for k in nodes if k != j]
        denom = 1.0
        for k in others:
            denom *= (j - k)
        deg = len(others) - power          # degree of the elementary symmetric sum
        esum = sum(math.prod(c) for c in combinations(others, deg)) if deg > 0 else 1.0
        coef = (-1.0) ** deg * esum        # x^power coefficient of prod (x - k)
        total += math.factorial(power) * coef / denom * grid[idx]
    return total / ff ** power


def point_energies(coords, bonds, nocc, points, step, eps=1e-11, kind="ohno-klopman", solver="ccsd"):
    """Total energies at the field points: oracle SCF, then oracle CCSD (or FCI
    through ``solver``, a callable (core, G, huckel, nocc) -> energy)."""
    G, core, hk = ppp.build(coords, bonds, kind=kind)
    out = []
    for p in points:
        c, h = set_core(core, hk, coords, p, step)
        if solver == "ccsd":
            r = scf.rhf(c, G, h, nocc)
            F, eref = scf.mo_fock(r["C"], c, G, nocc)
            res = ccsd.solve(r["C"], F, G, nocc, eref, eps=eps, diis_scale=True)
            assert res["status"] == 1
            out.append(res["energy"][0])
        else:
            out.append(solver(c, G, h, nocc))
    return np.array(out)
