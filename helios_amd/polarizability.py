"""Finite-field (hyper)polarizability as one batch of CCSD solves.

The reference's polarizability task (getPolarizability,
src/modules/mod_property.f90:405-506) visits every point of a field grid,
applies setCore (:467-486: the core diagonal gains F . r_mu, F = p * step, r
the centred coordinates of readMoleculeInformation.f90:381-397; perturbate,
src/modules/mod_helios.f90:577-601, adds the same to the Hueckel matrix),
reruns SCF + r-ccsd (getEnergy) and then differentiates the energies on the
grid by Lagrange central differences (putPoint.f90, showPolarizability.f90,
mod_derivative.f90 deLagDeriv, mod_math.f90:505-560 LagrangeDerivative).

Here the points' SCFs run on the host and all points are solved together by
one hcc_batch_solve; the derivatives and units follow showPolarizability:
dipole in Debye, alpha, beta, gamma in atomic units.

    python -m helios_amd.polarizability            # octatetraene, x, 5 points (GPU)
"""
import argparse
import json
import math
import sys

import numpy as np

from helios_amd import capi, molecules, ppp

# src/modules/mod_helios.f90:42-43,50
BOHR_RADIUS = 0.52917721067     # Angstrom
HARTREE_ENERGY = 27.21138602    # eV
DIPOLE_TO_DEBYE = 4.8031


def unit_factor(power):
    """showPolarizability's transition(power): -Debye per e*A for the first
    derivative, -Eh^(k-1) / a0^k for the k-th (energies in eV, fields in V/A)."""
    if power == 1:
        return -DIPOLE_TO_DEBYE
    return -HARTREE_ENERGY ** (power - 1) / BOHR_RADIUS ** power


AXES = "xyz"


def field_points(scales="x", npoints=5):
    """Grid points (px, py, pz) in the reference's order for the symmetry-free
    case (src/general/symmertySettings.f90:30-140): the zero point first, then
    the axis grid (one axis), the full plane (two axes), or the three axes and
    their planes without the fully off-axis points (three axes)."""
    if npoints < 1 or npoints % 2 == 0 or npoints > 9:
        raise ValueError("points must be odd, 1..9")
    if scales not in ("x", "y", "z", "xy", "yx", "xz", "zx", "yz", "zy", "xyz"):   # definebd.f90:75
        raise ValueError(f"unsupported components {scales!r}")
    axes = sorted({AXES.index(c) for c in scales})
    sto = (npoints - 1) // 2
    rng = range(-sto, sto + 1)
    pts = [(0, 0, 0)]
    if len(axes) == 1:
        for i in rng:
            if i:
                p = [0, 0, 0]
                p[axes[0]] = i
                pts.append(tuple(p))
    elif len(axes) == 2:
        for i in rng:
            for j in rng:
                if i or j:
                    p = [0, 0, 0]
                    p[axes[0]], p[axes[1]] = i, j
                    pts.append(tuple(p))
    else:
        for i in rng:
            for j in rng:
                for k in rng:
                    if (i and j and k) or not (i or j or k):
                        continue
                    pts.append((i, j, k))
    return pts


def centred(coords):
    x = np.asarray(coords, dtype=np.float64)
    return x - x.mean(0)


def core_perturbation(coords, point, step):
    """setCore: diagonal F . r_mu with F = point * step (eV for r in A, F in V/A)."""
    r = centred(coords)
    return np.diag(r @ (np.asarray(point, dtype=np.float64) * step))


def lagrange_weights(npoints, power):
    """Weights w_j (j = -s..s) of the power-th derivative at 0 of the
    polynomial interpolating f at the integer nodes -s..s: the solution of
    sum_j w_j j^m = power! delta_{m,power}, m = 0..npoints-1."""
    s = (npoints - 1) // 2
    nodes = np.arange(-s, s + 1, dtype=np.float64)
    V = np.vander(nodes, npoints, increasing=True).T
    rhs = np.zeros(npoints)
    rhs[power] = math.factorial(power)
    return np.linalg.solve(V, rhs)


def derivative(grid, spec, npoints, step):
    """deLagDeriv: d^|spec| E / d F_spec at zero field from the energy grid
    grid[(px, py, pz)]; spec like 'x', 'xx', 'xxyy'.  Mixed derivatives take the
    inner axis first (mod_derivative.f90:88-96)."""
    counts = [spec.count(c) for c in AXES]
    if sum(counts) > npoints - 1:
        return None
    s = (npoints - 1) // 2
    used = [a for a in range(3) if counts[a]]

    def line(axis, power, fixed):
        w = lagrange_weights(npoints, power)
        tot = 0.0
        for q, j in enumerate(range(-s, s + 1)):
            p = list(fixed)
            p[axis] = j
            tot += w[q] * grid[tuple(p)]
        return tot / step ** power

    if len(used) == 1:
        return line(used[0], counts[used[0]], (0, 0, 0))
    if len(used) == 2:
        a, b = used
        inner = {}
        for i in range(-s, s + 1):
            p = [0, 0, 0]
            p[a] = i
            inner[i] = line(b, counts[b], p)
        w = lagrange_weights(npoints, counts[a])
        return sum(w[q] * inner[j] for q, j in enumerate(range(-s, s + 1))) / step ** counts[a]
    raise ValueError("three-axis mixed derivatives are not supported")


def prepare_points(mol, points, step, kind="ohno-klopman"):
    """SCF of every field point (host): solver inputs for the batch."""
    G, core, hk = ppp.model(mol.coords, mol.bonds, mol.alpha, mol.U, 1, kind)
    out = []
    for p in points:
        P = core_perturbation(mol.coords, p, step)
        r = ppp.rhf(core + P, G, hk + P, mol.nocc)
        out.append(dict(C=r["C"], F=r["Fmo"], G=G, nocc=mol.nocc, eref=r["eref"], point=p))
    return out


def rotate_amplitudes(t1, t2, C_from, C_to, o):
    """Amplitudes of one field point carried into the orbitals of another,
    with U = C_to^T C_from (the site basis is orthonormal): t2'_ijab =
    Uo_ik Uo_jl Uv_ac Uv_bd t2_klcd and t1' = Uo t1 Uv^T + X^T, so that a sign
    flip or a rotation among near-degenerate orbitals between the points
    carries over.  X = U_vo Uo^-1 is the occupied-virtual mixing between the
    two references written as a singles excitation (Thouless: the old
    reference is e^X times the new one, up to normalisation); its other
    first-order effects are left out (a guess only)."""
    U = C_to.T @ C_from
    Uo, Uv = U[:o, :o], U[o:, o:]
    X = U[o:, :o] @ np.linalg.inv(Uo)
    return (Uo @ t1 @ Uv.T + X.T,
            np.einsum("ik,jl,klcd,ac,bd->ijab", Uo, Uo, t2, Uv, Uv, optimize=True))


def compute(mol, scales="x", npoints=5, step=0.02, precision=capi.HCC_FP64, eps=1e-10, diis_steps=20,
            guess=capi.GUESS_MP2, maxiter=500, warm=False):
    """Energies at every field point (one batched GPU solve) and their
    derivatives.  Returns dict(points, energy, status, iterations, dipole,
    alpha, beta, gamma, gamma_mixed) with the reference's units.

    warm: solve the zero-field point first and start every point from its
    amplitudes (rotated into the point's orbitals, HCC_GUESS_CURRENT) instead
    of from ``guess``; the reference starts every point from guessCC
    (mod_property.f90:467-486 calls getEnergy per point)."""
    points = field_points(scales, npoints)
    inputs = prepare_points(mol, points, step)
    ctx = capi.Context(0, precision, diis_steps)
    ctx.batch_load(inputs)
    if warm:
        o, v = mol.nocc, mol.n - mol.nocc
        z = points.index((0, 0, 0))
        zero = capi.Context(0, precision, diis_steps)
        zero.batch_load([inputs[z]])
        zero.batch_solve(guess, maxiter, eps)
        t1, t2 = zero.batch_amplitudes(0, o, v)
        zero.close()
        for k, inp in enumerate(inputs):
            ctx.batch_set_amplitudes(k, *rotate_amplitudes(t1, t2, inputs[z]["C"], inp["C"], o))
        ctx.batch_solve(capi.GUESS_CURRENT, maxiter, eps)
    else:
        ctx.batch_solve(guess, maxiter, eps)
    e, it, st = ctx.batch_results()
    ctx.close()
    grid = {p: e[k, 0] for k, p in enumerate(points)}
    axes = sorted({c for c in scales})
    res = dict(points=points, energy=e[:, 0].copy(), status=st, iterations=it, dipole={}, alpha={}, beta={},
               gamma={}, gamma_mixed={})
    for a in axes:
        res["dipole"][a] = derivative(grid, a, npoints, step) * unit_factor(1)
        res["alpha"][a + a] = derivative(grid, a * 2, npoints, step) * unit_factor(2)
        if npoints >= 5:
            res["beta"][a * 3] = derivative(grid, a * 3, npoints, step) * unit_factor(3)
            res["gamma"][a * 4] = derivative(grid, a * 4, npoints, step) * unit_factor(4)
    if len(axes) == 2 and npoints >= 5:
        a, b = axes
        res["gamma_mixed"][a * 2 + b * 2] = derivative(grid, a * 2 + b * 2, npoints, step) * unit_factor(4)
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--molecule", default="octatetraene", help="octatetraene | polyeneN | fragN:seed")
    ap.add_argument("--components", default="x")
    ap.add_argument("--points", type=int, default=5)
    ap.add_argument("--step", type=float, default=0.02)
    ap.add_argument("--mixed", action="store_true", help="mixed-precision solver")
    ap.add_argument("--eps", type=float, default=1e-10)
    ap.add_argument("--warm", action="store_true", help="start every point from the zero-field amplitudes")
    a = ap.parse_args()
    if a.molecule == "octatetraene":
        mol = molecules.octatetraene()
    elif a.molecule.startswith("polyene"):
        mol = molecules.polyene(int(a.molecule[7:]))
    elif a.molecule.startswith("frag"):
        n, seed = a.molecule[4:].split(":")
        mol = molecules.random_fragment(int(n), seed=int(seed))
    else:
        sys.exit(f"unknown molecule {a.molecule}")
    r = compute(mol, a.components, a.points, a.step, capi.HCC_MIXED if a.mixed else capi.HCC_FP64, a.eps,
                warm=a.warm)
    out = {k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in r.items()}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
