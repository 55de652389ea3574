"""Finite-field polarizability pieces that need no GPU.

* field grids: the reference's point sets for one, two and three components;
* Lagrange derivatives: the product's weights (a Vandermonde solve) against
  the oracle's restatement of the reference's elementary-symmetric-product
  construction, and both exact on polynomials of degree below the number of
  points (a wrong sign, node or factorial fails this);
* mixed derivatives on an analytic two-field energy surface;
* physics pin: for two electrons CCSD is exact, so at every field point the
  oracle CCSD energy equals the determinant FCI energy, and so do the
  polarizabilities derived from them.
"""
import math

import numpy as np
import pytest

from helios_amd import molecules as M, polarizability as pol
from oracle import ccsd, fci, field, ppp, scf


def test_field_point_sets():
    assert pol.field_points("x", 5) == [(0, 0, 0), (-2, 0, 0), (-1, 0, 0), (1, 0, 0), (2, 0, 0)]
    assert len(pol.field_points("xy", 5)) == 25 and len(set(pol.field_points("xy", 5))) == 25
    pts = pol.field_points("xyz", 3)
    assert len(pts) == 1 + 3 * 2 + 3 * 4     # axes and planes, no point with all three nonzero
    assert all(not (i and j and k) for i, j, k in pts)
    with pytest.raises(ValueError):
        pol.field_points("x", 4)


@pytest.mark.parametrize("npoints", [3, 5, 7, 9])
def test_lagrange_weights(npoints):
    s = (npoints - 1) // 2
    for power in range(1, npoints):
        w = pol.lagrange_weights(npoints, power)
        # the oracle's construction from the reference's formula
        grid_e = lambda k: np.eye(npoints)[k]          # noqa: E731
        wo = np.array([field.lagrange_derivative(npoints, power, grid_e(k), 1.0) for k in range(npoints)])
        assert np.abs(w - wo).max() < 1e-9 * max(1.0, np.abs(w).max())
        # exact for every monomial of degree < npoints (a spacing exact in
        # binary, so that only the weights can err)
        for m in range(npoints):
            h = 0.5
            vals = np.array([(j * h) ** m for j in range(-s, s + 1)])
            d = field.lagrange_derivative(npoints, power, vals, h)
            exact = math.factorial(m) if m == power else 0.0
            assert abs(d - exact) < 1e-6 * max(1.0, abs(exact))


def test_mixed_derivative_on_analytic_surface():
    # E(F) = a Fx + b Fy + c Fx^2 + d Fy^2 + e Fx^2 Fy^2 + f Fx^4
    a, b, c, d, e, f = 0.3, -0.2, -1.7, -2.9, 5.5, 0.7
    h = 0.02
    pts = pol.field_points("xy", 5)
    grid = {}
    for p in pts:
        x, y = p[0] * h, p[1] * h
        grid[p] = a * x + b * y + c * x * x + d * y * y + e * x * x * y * y + f * x ** 4
    assert abs(pol.derivative(grid, "x", 5, h) - a) < 1e-9
    assert abs(pol.derivative(grid, "yy", 5, h) - 2 * d) < 1e-7
    assert abs(pol.derivative(grid, "xxyy", 5, h) - 4 * e) < 1e-4
    assert abs(pol.derivative(grid, "xxxx", 5, h) - 24 * f) < 1e-3
    assert pol.derivative(grid, "xxxxx", 5, h) is None


def test_two_electron_ccsd_equals_fci_in_a_field():
    mol = M.ethylene(1.35)
    pts = pol.field_points("x", 5)
    fci_solver = lambda c, G, h, nocc: fci.fci_energy(scf.rhf(c, G, h, nocc)["C"], G, c, nocc)   # noqa: E731
    e_cc = field.point_energies(mol.coords, mol.bonds, 1, pts, 0.02)
    e_fci = field.point_energies(mol.coords, mol.bonds, 1, pts, 0.02, solver=fci_solver)
    assert np.abs(e_cc - e_fci).max() < 1e-9
    order = [pts.index((j, 0, 0)) for j in range(-2, 3)]
    a_cc = field.lagrange_derivative(5, 2, e_cc[order], 0.02)
    a_fci = field.lagrange_derivative(5, 2, e_fci[order], 0.02)
    assert a_cc < 0.0 and abs(a_cc - a_fci) < 1e-5 * abs(a_fci)


def test_product_field_inputs_match_oracle_scf():
    """The product's per-point SCF inputs give the oracle's CCSD energies."""
    mol = M.octatetraene()
    pts = pol.field_points("x", 3)
    ins = pol.prepare_points(mol, pts, 0.02)
    e_or = field.point_energies(mol.coords, mol.bonds, 4, pts, 0.02, eps=1e-10)
    for k, p in enumerate(ins):
        r = ccsd.solve(p["C"], p["F"], p["G"], 4, p["eref"], eps=1e-10, diis_scale=True)
        assert abs(r["energy"][0] - e_or[k]) < 1e-8


def test_rotated_amplitudes_solve_the_rotated_equations():
    """CCSD is invariant under rotations among the occupied and among the
    virtual orbitals: the amplitudes solved in rotated orbitals C' = C Q
    (Q block-orthogonal) equal the ones solved in C and carried over by
    rotate_amplitudes.  A transposed block or a missing factor fails."""
    from tests._molprep import prepare
    mol = M.octatetraene()
    p = prepare(mol)
    o, n = mol.nocc, p["n"]
    r0 = ccsd.solve(p["C"], p["F"], p["G"], o, p["eref"], eps=1e-11, diis_scale=True)
    rng = np.random.default_rng(4)
    Q = np.zeros((n, n))
    Q[:o, :o] = np.linalg.qr(rng.standard_normal((o, o)))[0]
    Q[o:, o:] = np.linalg.qr(rng.standard_normal((n - o, n - o)))[0]
    C1 = p["C"] @ Q
    F1 = Q.T @ p["F"] @ Q
    r1 = ccsd.solve(C1, F1, p["G"], o, p["eref"], eps=1e-11, diis_scale=True)
    t1, t2 = pol.rotate_amplitudes(r0["t1"], r0["t2"], p["C"], C1, o)
    assert abs(r1["energy"][0] - r0["energy"][0]) < 1e-9
    assert np.abs(t1 - r1["t1"]).max() < 1e-8 and np.abs(t2 - r1["t2"]).max() < 1e-8
