"""Singlet LR-CCSD on the GPU (hcc_batch_lr_sigma + helios_amd.lrccsd)
against the oracle (oracle/lrccsd.py, pinned by the two-electron full CI):

* the device's sigma = J r at random pair-symmetric r equals the oracle's
  central difference of the same residual (1e-6 of max |sigma|), for every
  molecule of a batch, and the amplitudes are left as they were;
* the three lowest singlet excitation energies of a batch (polyenes with
  their doubly excited lowest state, honeycomb fragments) equal the lowest
  real eigenvalues of the oracle's dense Jacobian within 2e-6 eV, from unit
  singles and from the CIS start vectors, and their average excitation
  levels and reference weights r0 equal those of the dense eigenvectors;
* for two electrons they equal the singlet full-CI excitation energies;
* a molecule's roots in the batch equal its roots alone;
* the device vector space's operations (dot products, max-norms, linear
  combinations, Gram blocks, block combinations, the orbital-energy
  preconditioner, sigma between slots) equal their numpy counterparts, and
  invalid jobs are rejected.
"""
import numpy as np
import pytest

from helios_amd import capi, lrccsd, molecules as M
from oracle import ccsd, lrccsd as olr, ppp, scf
from tests._molprep import prepare

pytestmark = pytest.mark.gpu


def _solved_ctx(preps):
    ctx = capi.Context(0, capi.HCC_FP64, 20)
    ctx.batch_load(preps)
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-10)
    _, _, st = ctx.batch_results()
    assert (st == 1).all()
    return ctx


def _pair_symmetric(rng, o, v):
    r2 = rng.standard_normal((o, o, v, v))
    return 0.5 * (r2 + r2.transpose(1, 0, 3, 2))


def test_sigma_matches_oracle():
    mols = [M.octatetraene(), M.random_fragment(12, seed=2), M.batch_member(11)]
    preps = [prepare(m) for m in mols]
    ctx = _solved_ctx(preps)
    rng = np.random.default_rng(5)
    r1s, r2s = [], []
    for m in mols:
        o, v = m.nocc, m.n - m.nocc
        r1s.append(rng.standard_normal((o, v)))
        r2s.append(_pair_symmetric(rng, o, v))
    before = ctx.batch_results()[0]
    s1, s2 = ctx.batch_lr_sigma(np.concatenate([x.ravel() for x in r1s]),
                                np.concatenate([x.ravel() for x in r2s]), 1e-5)
    a1 = a2 = 0
    for m, p, r1, r2 in zip(mols, preps, r1s, r2s):
        o, v = m.nocc, m.n - m.nocc
        ref = ccsd.solve(p["C"], p["F"], p["G"], o, p["eref"], eps=1e-11, diis_scale=True)
        q1, q2 = olr.sigma(p["C"], p["F"], p["G"], o, ref["t1"], ref["t2"], r1, r2, 1e-5)
        g1 = s1[a1:a1 + o * v].reshape(o, v)
        g2 = s2[a2:a2 + o * o * v * v].reshape(o, o, v, v)
        scale = max(np.abs(q1).max(), np.abs(q2).max())
        print(m.name, "max|dsigma|", np.abs(g1 - q1).max(), np.abs(g2 - q2).max(), "of", scale)
        assert np.abs(g1 - q1).max() < 1e-6 * scale and np.abs(g2 - q2).max() < 1e-6 * scale
        a1, a2 = a1 + o * v, a2 + o * o * v * v
    # amplitudes restored: a second sigma gives the same numbers, energies unchanged
    t1, t2 = ctx.batch_lr_sigma(np.concatenate([x.ravel() for x in r1s]), np.concatenate([x.ravel() for x in r2s]),
                                1e-5)
    assert np.array_equal(t1, s1) and np.array_equal(t2, s2)
    assert np.array_equal(ctx.batch_results()[0], before)
    ctx.close()


@pytest.mark.parametrize("guess", ["unt", "cis"])
def test_lowest_roots_match_dense_jacobian(guess):
    mols = [M.octatetraene(), M.polyene(6), M.random_fragment(10, seed=4), M.random_fragment(12, seed=2)]
    preps = [prepare(m) for m in mols]
    ctx = _solved_ctx(preps)
    gammas = [p["G"] for p in preps] if guess == "cis" else None
    om, res, its, conv, ael, r0 = lrccsd.excitation_energies(ctx, [(m.nocc, m.n - m.nocc) for m in mols],
                                                             nroots=3, gammas=gammas)
    ctx.close()
    for m, p, w, c, it, lv, q in zip(mols, preps, om, conv, its, ael, r0):
        r = ccsd.solve(p["C"], p["F"], p["G"], m.nocc, p["eref"], eps=1e-11, diis_scale=True)
        lam, _, vec = olr.excitation_energies(p["C"], p["F"], p["G"], m.nocc, r["t1"], r["t2"], vectors=True)
        ref_ael = [olr.average_excitation_level(*x) for x in vec[:3]]
        print(m.name, "gpu", w, "oracle", lam[:3], "iterations", it, "diff", np.abs(w - lam[:3]).max(),
              "AEL", lv, ref_ael, "r0", q)
        assert c
        assert np.abs(w - lam[:3]).max() < 2e-6
        assert np.abs(lv - ref_ael).max() < 1e-4
        if guess == "cis":
            want = [olr.reference_weight(p["C"], p["F"], p["G"], m.nocc, r["t1"], lam[k], *vec[k]) for k in range(3)]
            assert np.abs(q - want).max() < 1e-5


def test_two_electrons_equal_fci():
    mols = [M.ethylene(), M.polyene(4), M.random_fragment(6, seed=3)]
    preps, fcis = [], []
    for m in mols:
        G, core, hk = ppp.build(m.coords, m.bonds, alpha=m.alpha, U=m.U, kind="ohno-klopman")
        r = scf.rhf(core, G, hk, 1)
        F, eref = scf.mo_fock(r["C"], core, G, 1)
        preps.append(dict(C=r["C"], F=F, G=G, nocc=1, eref=eref, n=m.n))
        fcis.append(olr.singlets_two_electrons(r["C"], core, G))
    ctx = _solved_ctx(preps)
    om, res, its, conv, _, _ = lrccsd.excitation_energies(ctx, [(1, m.n - 1) for m in mols], nroots=2)
    ctx.close()
    for m, w, f, c in zip(mols, om, fcis, conv):
        print(m.name, "gpu", w, "fci", f[:2])
        assert c
        assert np.abs(w - f[:2]).max() < 2e-6


def test_batch_equals_alone():
    mols = [M.polyene(6), M.random_fragment(10, seed=4)]
    preps = [prepare(m) for m in mols]
    ctx = _solved_ctx(preps)
    om, _, _, _, _, _ = lrccsd.excitation_energies(ctx, [(m.nocc, m.n - m.nocc) for m in mols], nroots=2)
    ctx.close()
    for k, (m, p) in enumerate(zip(mols, preps)):
        one = _solved_ctx([p])
        w1, _, _, _, _, _ = lrccsd.excitation_energies(one, [(m.nocc, m.n - m.nocc)], nroots=2)
        one.close()
        assert np.abs(w1[0] - om[k]).max() < 1e-6


def test_vector_space_operations():
    mols = [M.octatetraene(), M.random_fragment(12, seed=2), M.batch_member(11)]
    preps = [prepare(m) for m in mols]
    ctx = _solved_ctx(preps)
    ctx.lr_space(6)
    rng = np.random.default_rng(7)
    X = []
    for k, m in enumerate(mols):
        o, v = m.nocc, m.n - m.nocc
        rows = []
        for sl in range(6):
            r1 = rng.standard_normal(o * v)
            r2 = _pair_symmetric(rng, o, v).ravel()
            x = np.concatenate([r1, r2])
            ctx.lr_put(k, sl, x)
            rows.append(x)
        X.append(np.array(rows))
    for k in range(len(mols)):
        assert np.array_equal(ctx.lr_get(k, 3, X[k].shape[1]), X[k][3])
    jobs = [(k, a, b) for k in range(len(mols)) for a in range(3) for b in range(2, 5)]
    got = ctx.lr_dots(jobs)
    ref = np.array([X[k][a] @ X[k][b] for k, a, b in jobs])
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    got = ctx.lr_dots([(0, 1, 2), (1, 4, -1), (2, 0, 5)], mode=1)
    ref = [np.sqrt(X[0][1] ** 2 + X[0][2] ** 2).max(), np.abs(X[1][4]).max(), np.sqrt(X[2][0] ** 2 + X[2][5] ** 2).max()]
    assert np.allclose(got, ref, rtol=1e-15, atol=0)
    # dst = beta dst + sum coef slot
    ctx.lr_combine(np.array([[0, 5, 0, 2], [1, 0, 2, 3], [2, 4, 5, 0]]), np.array([0.5, 0.0, -2.0]),
                   np.array([1, 2, 1, 2, 3]), np.array([0.25, -1.5, 2.0, 3.0, -0.75]))
    want = {(0, 5): 0.5 * X[0][5] + 0.25 * X[0][1] - 1.5 * X[0][2],
            (1, 0): 2.0 * X[1][1] + 3.0 * X[1][2] - 0.75 * X[1][3],
            (2, 4): -2.0 * X[2][4]}
    for (k, sl), w in want.items():
        g = ctx.lr_get(k, sl, X[k].shape[1])
        assert np.abs(g - w).max() < 1e-13 * np.abs(w).max()
        X[k][sl] = g
    # orbital-energy preconditioner
    ctx.lr_precondition(np.array([[0, 1], [2, 3]]), np.array([3.9, 4.4]), 1e-2)
    for (k, sl), w in [((0, 1), 3.9), ((2, 3), 4.4)]:
        m = mols[k]
        o = m.nocc
        _, F, _ = ctx.batch_orbitals(k, m.n)
        e = np.diag(F)
        D = np.concatenate([(e[o:][None, :] - e[:o][:, None]).ravel(),
                            (e[o:][None, None, :, None] + e[o:][None, None, None, :] - e[:o][:, None, None, None]
                             - e[:o][None, :, None, None]).ravel()])
        den = w - D
        den = np.where(np.abs(den) < 1e-2, np.copysign(1e-2, den), den)
        g = ctx.lr_get(k, sl, X[k].shape[1])
        assert np.abs(g - X[k][sl] / den).max() < 1e-14 * np.abs(g).max()
        X[k][sl] = g
    # sigma between slots = sigma of the host vectors, molecule 1 skipped
    ctx.lr_sigma_slots(np.array([2, -1, 4]), np.array([0, 0, 1]), 1e-5)
    r1 = np.concatenate([X[k][sl][:mols[k].nocc * (mols[k].n - mols[k].nocc)] for k, sl in enumerate([2, 2, 4])])
    r2 = np.concatenate([X[k][sl][mols[k].nocc * (mols[k].n - mols[k].nocc):] for k, sl in enumerate([2, 2, 4])])
    s1, s2 = ctx.batch_lr_sigma(r1, r2, 1e-5)
    a1 = a2 = 0
    for k, m in enumerate(mols):
        o, v = m.nocc, m.n - m.nocc
        g = ctx.lr_get(k, [0, 0, 1][k], X[k].shape[1])
        if k == 1:
            assert np.array_equal(g, X[k][0])        # skipped: slot untouched
        else:
            assert np.array_equal(g, np.concatenate([s1[a1:a1 + o * v], s2[a2:a2 + o * o * v * v]]))
        a1, a2 = a1 + o * v, a2 + o * o * v * v
    # invalid jobs are rejected before any work
    with pytest.raises(capi.HccError):
        ctx.lr_combine(np.array([[0, 1, 0, 1]]), np.array([0.0]), np.array([1]), np.array([1.0]))   # dst is a term
    with pytest.raises(capi.HccError):
        ctx.lr_dots([(0, 6, 0)])
    with pytest.raises(capi.HccError):
        ctx.lr_sigma_slots(np.array([0, -1, 0]), np.array([0, 0, 1]), 1e-5)   # in == out
    ctx.close()


def test_block_gram_and_combinations():
    """hcc_lr_gram and hcc_lr_combine_block against numpy on the same slots:
    blocks wider than one 16 x 64 piece, a single pair, vector lengths that
    end inside a 16-element run, a 64-element step and a 4096-element chunk
    (N = 272, 2450, 50850), slot lists out of order, beta on and off, a
    scaling with no terms; invalid jobs are rejected before any work."""
    mols = [M.octatetraene(), M.random_fragment(14, seed=3), M.random_fragment(30, seed=5)]
    preps = [prepare(m) for m in mols]
    ctx = _solved_ctx(preps)
    nsl = 96
    ctx.lr_space(nsl)
    rng = np.random.default_rng(11)
    X = []
    for k, m in enumerate(mols):
        n = m.nocc * (m.n - m.nocc) * (1 + m.nocc * (m.n - m.nocc))
        rows = rng.standard_normal((nsl, n))
        for sl in range(nsl):
            ctx.lr_put(k, sl, rows[sl])
        X.append(rows)
    perm = rng.permutation(nsl)
    blocks = [(0, perm[:20], perm[20:90]), (1, perm[5:6], perm[7:8]), (2, perm[:20], perm[20:90]),
              (2, perm[90:93], perm[90:93]), (1, perm[:17], perm[30:95])]
    heads, sl = [], []
    at = 0
    for k, a, b in blocks:
        heads.append((k, len(a), len(b), at, at + len(a)))
        sl.extend(a)
        sl.extend(b)
        at += len(a) + len(b)
    g = ctx.lr_gram(np.array(heads), np.array(sl))
    at = 0
    for k, a, b in blocks:
        want = X[k][a] @ X[k][b].T
        bound = np.abs(X[k][a]) @ np.abs(X[k][b]).T
        got = g[at:at + want.size].reshape(want.shape)
        assert np.all(np.abs(got - want) <= 1e-13 * bound)
        at += want.size
    # block combinations
    jobs = [(0, perm[:18], 0.5, perm[18:88]), (2, perm[:18], 0.0, perm[18:88]), (1, perm[40:41], 2.0, []),
            (2, perm[88:91], -1.0, perm[91:96])]
    heads, sl, cf, beta = [], [], [], []
    at = nc = 0
    coefs = []
    for k, d, bt, t in jobs:
        C = rng.standard_normal((len(d), len(t)))
        coefs.append(C)
        heads.append((k, len(d), len(t), at, at + len(d), nc))
        sl.extend(d)
        sl.extend(t)
        cf.extend(C.ravel())
        beta.append(bt)
        at += len(d) + len(t)
        nc += C.size
    ctx.lr_combine_block(np.array(heads), np.array(beta), np.array(sl), np.array(cf))
    for (k, d, bt, t), C in zip(jobs, coefs):
        want = bt * X[k][d] + (C @ X[k][t] if len(t) else 0.0)
        bound = abs(bt) * np.abs(X[k][d]) + (np.abs(C) @ np.abs(X[k][t]) if len(t) else 0.0)
        for q, s in enumerate(d):
            got = ctx.lr_get(k, s, X[k].shape[1])
            assert np.all(np.abs(got - want[q]) <= 1e-13 * bound[q])
            X[k][s] = got
    # untouched slots stay as they were
    for k in range(3):
        assert np.array_equal(ctx.lr_get(k, perm[95], X[k].shape[1]), X[k][perm[95]])
    with pytest.raises(capi.HccError):   # destination among the terms
        ctx.lr_combine_block(np.array([[0, 1, 1, 0, 1, 0]]), np.array([0.0]), np.array([3, 3]), np.array([1.0]))
    with pytest.raises(capi.HccError):   # repeated destination
        ctx.lr_combine_block(np.array([[0, 2, 1, 0, 2, 0]]), np.array([0.0]), np.array([3, 3, 4]),
                             np.array([1.0, 1.0]))
    with pytest.raises(capi.HccError):   # slot-list range past its end
        ctx.lr_gram(np.array([[0, 2, 2, 0, 1]]), np.array([1, 2]))
    with pytest.raises(capi.HccError):   # slot out of range
        ctx.lr_gram(np.array([[0, 1, 1, 0, 1]]), np.array([1, nsl]))
    ctx.close()


def test_replicas_equal_single_copies():
    """Three copies of each molecule, loaded from the same inputs and solved
    together: with the copies applying J to three vectors per call, the roots,
    residuals, iteration counts, excitation levels and r0 equal those of the
    molecules alone bit for bit (the copies' amplitudes and products are the
    molecule's own); invalid sources are rejected."""
    mols = [M.polyene(6), M.random_fragment(10, seed=4), M.random_fragment(12, seed=2)]
    preps = [prepare(m) for m in mols]
    dims = [(m.nocc, m.n - m.nocc) for m in mols]
    gam = [p["G"] for p in preps]
    ctx = _solved_ctx(preps)
    ref = lrccsd.excitation_energies(ctx, dims, nroots=3, gammas=gam)
    ctx.close()
    R = 3
    rep = _solved_ctx([p for p in preps for _ in range(R)])
    got = lrccsd.excitation_energies(rep, dims, nroots=3, gammas=gam,
                                     replicas=[[R * k + j for j in range(R)] for k in range(len(mols))])
    print("iterations", ref[2], got[2])
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    nm = R * len(mols)
    src = np.zeros(nm, dtype=np.int32)
    ins = np.full(nm, -1, dtype=np.int32)
    outs = np.zeros(nm, dtype=np.int32)
    ins[R] = 0
    outs[R] = 1
    with pytest.raises(capi.HccError):   # a copy of the second molecule takes the third's space as source
        src[R] = R * 2
        rep.lr_sigma_from(src, ins, outs, 1e-5)
    with pytest.raises(capi.HccError):   # two copies write the same slot
        src[:] = 0
        ins[:2] = [0, 2]
        outs[:2] = [1, 1]
        ins[R] = -1
        rep.lr_sigma_from(src, ins, outs, 1e-5)
    with pytest.raises(capi.HccError):   # a copy without a space
        rep.lr_put(1, 0, np.zeros(dims[0][0] * dims[0][1] * (1 + dims[0][0] * dims[0][1])))
    rep.close()


def test_device_davidson_equals_oracle_driven_at_n30():
    """Beyond the sizes whose dense Jacobian can be formed: for a 30-site
    batch member (25650 coordinates) the device Davidson (vectors in device
    memory, sigma from the GPU residual) and the same algorithm driven by the
    oracle (numpy vectors, the oracle's residual) reach the same three roots,
    average excitation levels and reference weights."""
    from tests.test_lrccsd_oracle import _OracleCtx
    mol = M.batch_member(13)
    p = prepare(mol)
    o, v = mol.nocc, mol.n - mol.nocc
    ctx = _solved_ctx([p])
    om, res, its, conv, ael, r0 = lrccsd.excitation_energies(ctx, [(o, v)], nroots=3, gammas=[p["G"]])
    ctx.close()
    ref = ccsd.solve(p["C"], p["F"], p["G"], o, p["eref"], eps=1e-11, diis_scale=True)
    om2, res2, its2, conv2, ael2, r02 = lrccsd.excitation_energies(
        _OracleCtx([(p, o, ref["t1"], ref["t2"])]), [(o, v)], nroots=3, gammas=[p["G"]])
    print("gpu", om[0], its[0], "oracle-driven", om2[0], its2[0], "AEL", ael[0], ael2[0], "r0", r0[0], r02[0])
    assert conv[0] and conv2[0]
    assert np.abs(om[0] - om2[0]).max() < 2e-6
    assert np.abs(ael[0] - ael2[0]).max() < 1e-4 and np.abs(r0[0] - r02[0]).max() < 1e-4
