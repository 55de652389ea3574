"""Seeded synthetic planar carbon pi-systems for the r-ccsd path.

Every site is carbon with one pi electron; bonds are the pairs closer than
1.6 A, with resonance integral beta = -2.4 eV and kind ``nonset``; molecules
lie in the z=0 plane.  These are the inputs the reference reads from its
``molecule`` block (layout: src/input/definebdDescription.f90:218-242).

Generators:
  * ``polyene(n)``         all-trans zigzag chain, C=C 1.35 A / C-C 1.46 A, 120 deg
  * ``honeycomb_flake(n)`` the n honeycomb sites nearest a hexagon centre,
                           1.40 A bonds, seeded +-0.02 A jitter
  * ``random_fragment``    seeded random accretion on a 1.40 A honeycomb,
                           rejecting sublattice-imbalanced or small-gap
                           (Hueckel gap < 0.5 eV) fragments
"""
import math
from dataclasses import dataclass, field

import numpy as np

BOND_CUTOFF = 1.6     # A
BETA = -2.4           # eV
U_CARBON = 11.13      # eV
ALPHA_CARBON = 0.0    # eV; the correlation energy does not depend on alpha


@dataclass
class Molecule:
    name: str
    coords: np.ndarray                       # (n, 3) Angstrom
    bonds: list = field(default_factory=list)  # (i, j, beta) 0-based
    alpha: float = ALPHA_CARBON
    U: float = U_CARBON
    rejected: int = 0                        # generator rejections before this one

    @property
    def n(self):
        return len(self.coords)

    @property
    def nocc(self):
        return self.n // 2


def bonds_by_distance(coords, cutoff=BOND_CUTOFF, beta=BETA):
    x = np.asarray(coords)
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    n = len(x)
    return [(i, j, beta) for i in range(n) for j in range(i + 1, n) if d[i, j] < cutoff]


def polyene(n, double=1.35, single=1.46):
    """All-trans zigzag chain; bond k (0-based) is double for even k."""
    pts = [np.zeros(3)]
    for k in range(n - 1):
        ang = math.radians(30.0 if k % 2 == 0 else -30.0)
        length = double if k % 2 == 0 else single
        pts.append(pts[-1] + length * np.array([math.cos(ang), math.sin(ang), 0.0]))
    coords = np.array(pts)
    return Molecule(f"polyene-{n}", coords, bonds_by_distance(coords))


def octatetraene():
    m = polyene(8)
    m.name = "octatetraene"
    return m


def ethylene(r=1.35):
    coords = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    return Molecule(f"ethylene-{r:.2f}", coords, bonds_by_distance(coords))


# ----------------------------------------------------------------------------
# honeycomb lattice
# ----------------------------------------------------------------------------
def _site_xy(n1, n2, s, d):
    """Sublattice s (0=A, 1=B) site of cell (n1, n2); origin at a hexagon centre."""
    a1 = np.array([math.sqrt(3.0) * d, 0.0])
    a2 = np.array([math.sqrt(3.0) * d / 2.0, 1.5 * d])
    base = n1 * a1 + n2 * a2
    # offsets put the centre of the hexagon A(0,0) B(0,0) A(0,1) B(1,0) A(1,0) B(1,-1) at the origin
    off = np.array([-math.sqrt(3.0) * d / 2.0, -d / 2.0]) if s == 0 else np.array([-math.sqrt(3.0) * d / 2.0, d / 2.0])
    return base + off


def _neighbours(site):
    n1, n2, s = site
    if s == 0:
        return [(n1, n2, 1), (n1, n2 - 1, 1), (n1 + 1, n2 - 1, 1)]
    return [(n1, n2, 0), (n1, n2 + 1, 0), (n1 - 1, n2 + 1, 0)]


def _lattice_check(d=1.40):
    """Neighbour table agrees with geometry (used by tests)."""
    for s in (0, 1):
        p = _site_xy(0, 0, s, d)
        for q in _neighbours((0, 0, s)):
            r = np.linalg.norm(_site_xy(*q, d) - p)
            if abs(r - d) > 1e-9:
                return False
    return True


def _jitter(coords2d, rng, amp):
    return coords2d + rng.uniform(-amp, amp, size=coords2d.shape)


def _to3d(xy):
    return np.concatenate([xy, np.zeros((len(xy), 1))], axis=1)


def honeycomb_flake(n=150, d=1.40, jitter=0.02, seed=2):
    """The n honeycomb sites nearest a hexagon centre, ties broken by polar
    angle, then +-jitter A uniform displacements of x and y."""
    R = int(math.sqrt(n)) + 4
    sites = [(n1, n2, s) for n1 in range(-R, R + 1) for n2 in range(-R, R + 1) for s in (0, 1)]
    xy = np.array([_site_xy(*t, d) for t in sites])
    r = np.round(np.sqrt((xy ** 2).sum(1)), 9)
    ang = np.round(np.mod(np.arctan2(xy[:, 1], xy[:, 0]), 2 * math.pi), 9)
    order = np.lexsort((ang, r))[:n]
    xy = xy[order]
    rng = np.random.default_rng(seed)
    xy = _jitter(xy, rng, jitter)
    coords = _to3d(xy)
    m = Molecule(f"flake-{n}", coords, bonds_by_distance(coords))
    m.sublattice = np.array([sites[k][2] for k in order])
    return m


def huckel_gap(n, bonds, beta=BETA):
    h = np.zeros((n, n))
    for i, j, b in bonds:
        h[i, j] = h[j, i] = b
    e = np.linalg.eigvalsh(h)
    return float(e[n // 2] - e[n // 2 - 1])


def random_fragment(n, seed=None, rng=None, d=1.40, jitter=0.02, min_gap=0.5, max_tries=10000):
    """Grow a fragment of n sites by random accretion from one hexagon; reject
    fragments with unequal sublattices or a Hueckel gap below min_gap."""
    if rng is None:
        rng = np.random.default_rng(seed)
    hexagon = [(0, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 1), (1, 0, 0), (1, -1, 1)]
    rejected = 0
    for _ in range(max_tries):
        frag = list(hexagon[:min(6, n)])
        fset = set(frag)
        while len(frag) < n:
            cand = sorted({q for p in frag for q in _neighbours(p) if q not in fset})
            q = cand[rng.integers(len(cand))]
            frag.append(q)
            fset.add(q)
        na = sum(1 for t in frag if t[2] == 0)
        xy = np.array([_site_xy(*t, d) for t in frag])
        bonds0 = [(i, j, BETA) for i in range(n) for j in range(i + 1, n) if frag[j] in _neighbours(frag[i])]
        if 2 * na != n or huckel_gap(n, bonds0) < min_gap:
            rejected += 1
            continue
        xy = _jitter(xy, rng, jitter)
        coords = _to3d(xy)
        m = Molecule(f"frag-{n}", coords, bonds_by_distance(coords), rejected=rejected)
        assert len(m.bonds) == len(bonds0)
        return m
    raise RuntimeError("random_fragment: no acceptable fragment")


def batch(count=4096, base_seed=3, nmin=10, nmax=48):
    """The batched workload: molecule k draws, from the generator seeded with
    base_seed + k, an even site count uniform in [nmin, nmax] and then its
    fragment."""
    return [batch_member(k, base_seed, nmin, nmax) for k in range(count)]


def batch_member(k, base_seed=3, nmin=10, nmax=48):
    rng = np.random.default_rng(base_seed + k)
    n = 2 * int(rng.integers(nmin // 2, nmax // 2 + 1))
    m = random_fragment(n, rng=rng)
    m.name = f"frag-{base_seed + k}-{n}"
    return m


def to_input(mol, method="r-ccsd", cc_accuracy=10, scf_accuracy=12, extra=""):
    """Input file text in the reference's block syntax (src/input/definebd.f90,
    molecule layout src/input/definebdDescription.f90:218-242)."""
    lines = [f"%general{{method={method}, jobtype=energy}}",
             "%parametrization{coulomb-integrals=ohno-klopman}",
             f"%scf{{accuracy={scf_accuracy}}}",
             f"%coupled-cluster{{accuracy={cc_accuracy}{', ' + extra if extra else ''}}}",
             "%molecule{", mol.name, str(mol.n), str(len(mol.bonds)), str(mol.n // 2), str(mol.n // 2)]
    for x, y, z in mol.coords:
        lines.append(f"C {x:.12f} {y:.12f} {z:.12f} {mol.alpha} {mol.U} 1")
    for i, j, b in mol.bonds:
        lines.append(f"{i + 1} {j + 1} {b} nonset")
    lines.append("}")
    return "\n".join(lines) + "\n"
