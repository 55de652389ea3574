"""Finite-field polarizability as one batch of GPU solves against the oracle.

The product (helios_amd.polarizability) runs the host SCF of every field
point and one hcc_batch_solve over all points; the oracle (oracle.field)
solves every point with its own SCF and CCSD.  Per-point energies within
1e-8 eV; dipole, alpha, beta and gamma from the same Lagrange derivatives.
For two electrons the GPU's alpha equals the determinant FCI's.
"""
import numpy as np
import pytest

from helios_amd import capi, molecules as M, polarizability as pol
from oracle import fci, field, scf

pytestmark = pytest.mark.gpu


def check(mol, scales, npoints, nocc, rel=1e-6):
    r = pol.compute(mol, scales, npoints, 0.02, eps=1e-11)
    e_or = field.point_energies(mol.coords, mol.bonds, nocc, r["points"], 0.02, eps=1e-11)
    de = np.abs(r["energy"] - e_or)
    print(mol.name, scales, "max |dE| per point", de.max(), "alpha", r["alpha"], "gamma", r["gamma"])
    assert (r["status"] == 1).all()
    assert de.max() < 1e-8
    grid = {p: e_or[k] for k, p in enumerate(r["points"])}
    for a, val in r["alpha"].items():
        ref = pol.derivative(grid, a, npoints, 0.02) * pol.unit_factor(2)
        assert abs(val - ref) < rel * abs(ref)
    for a, val in r["gamma"].items():
        ref = pol.derivative(grid, a, npoints, 0.02) * pol.unit_factor(4)
        assert abs(val - ref) < 1e-3 * abs(ref) + 1.0
    return r


def test_octatetraene_x_five_points():
    r = check(M.octatetraene(), "x", 5, 4)
    assert r["alpha"]["xx"] > 0.0


def test_fragment_xy_plane():
    r = check(M.random_fragment(10, seed=7), "xy", 5, 5)
    assert len(r["points"]) == 25 and "xxyy" in r["gamma_mixed"]


def test_two_electrons_alpha_equals_fci():
    mol = M.ethylene(1.40)
    r = pol.compute(mol, "x", 5, 0.02, eps=1e-12)
    fci_solver = lambda c, G, h, nocc: fci.fci_energy(scf.rhf(c, G, h, nocc)["C"], G, c, nocc)   # noqa: E731
    e_fci = field.point_energies(mol.coords, mol.bonds, 1, r["points"], 0.02, solver=fci_solver)
    grid = {p: e_fci[k] for k, p in enumerate(r["points"])}
    a_fci = pol.derivative(grid, "xx", 5, 0.02) * pol.unit_factor(2)
    print("ethylene alpha GPU", r["alpha"]["xx"], "FCI", a_fci)
    assert np.abs(r["energy"] - e_fci).max() < 1e-9
    assert abs(r["alpha"]["xx"] - a_fci) < 1e-6 * abs(a_fci)


def test_warm_start_from_the_zero_field_amplitudes():
    """Starting every field point from the zero-field amplitudes (rotated into
    its orbitals, HCC_GUESS_CURRENT) reaches the same energies in fewer
    iterations (a few per cent: the start is ~40x closer, but convergence to
    1e-10 is set by the asymptotic rate); HCC_GUESS_CURRENT without
    amplitudes for every molecule is refused."""
    from helios_amd import capi
    mol = M.random_fragment(10, seed=7)
    cold = pol.compute(mol, "xy", 5, 0.02, eps=1e-10)
    warm = pol.compute(mol, "xy", 5, 0.02, eps=1e-10, warm=True)
    print("iterations cold", cold["iterations"].mean(), "warm", warm["iterations"].mean())
    assert (warm["status"] == 1).all()
    assert np.abs(warm["energy"] - cold["energy"]).max() < 1e-9
    assert warm["iterations"].mean() < cold["iterations"].mean()
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    ctx.batch_load(pol.prepare_points(mol, pol.field_points("x", 3), 0.02))
    with pytest.raises(capi.HccError):
        ctx.batch_solve(capi.GUESS_CURRENT, 100, 1e-8)
    ctx.close()
