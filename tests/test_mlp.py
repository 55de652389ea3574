"""Pinning of the learned-guess oracle (oracle/mlp.py), CPU only.

* bf16(): hand-computed BF16 bit patterns (round to nearest, ties to even).
* backward(): every gradient element of the unrounded network against a
  central finite difference in float64 (a dropped term, a transposed weight
  or a wrong GELU derivative fails it).
* features(): a direct loop over the ZDO integrals (pq|rs) of oracle.ccsd.eri
  and the orbital energies, element by element; exact invariance under
  (i,a) <-> (j,b).
* adam(): the closed form of the first step, p - lr * g / (|g| + eps').
"""
import numpy as np
import pytest

from helios_amd import molecules as M
from oracle import ccsd, mlp
from tests._molprep import prepare


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7   # BF16 spacing in [1, 2)
    cases = {
        one: one,
        one + ulp / 2: one,                      # tie -> even (mantissa 0)
        one + 3 * ulp / 2: one + 2 * ulp,        # tie -> even (mantissa 2)
        one + ulp / 2 + 2.0 ** -20: one + ulp,   # just above the tie
        -(one + 3 * ulp / 2): -(one + 2 * ulp),
        3.0e38: 226 / 128 * 2.0 ** 127,          # 3e38 = 1.7632 * 2^127 -> 226/128 * 2^127
        1.0e-30: 1.0e-30,
    }
    for x, want in cases.items():
        got = mlp.bf16(np.float32(x))
        assert np.isclose(got, want, rtol=2 ** -9, atol=0), (x, got, want)
    # exactness of the tie cases
    assert mlp.bf16(np.float32(one + ulp / 2)) == one
    assert mlp.bf16(np.float32(one + 3 * ulp / 2)) == one + 2 * ulp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = mlp.init(3).astype(np.float64)
    p[mlp.OFF_W3:mlp.OFF_B3] *= 20.0          # a non-trivial output layer
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.normal(0, 0.3, 64)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.normal(0, 0.3, 64)
    X = rng.normal(0, 1, (24, mlp.NF))
    tg = rng.normal(0, 0.5, 24)
    wt = rng.uniform(0.5, 1.5, 24)
    _, g = mlp.backward(p, X, tg, wt, rounding=False)
    h = 1e-6
    fd = np.zeros_like(g)
    for k in range(mlp.NPARAM):
        pp = p.copy()
        pp[k] += h
        lp = mlp.loss(mlp.forward(pp, X, rounding=False)[0], tg, wt)
        pp[k] -= 2 * h
        lm = mlp.loss(mlp.forward(pp, X, rounding=False)[0], tg, wt)
        fd[k] = (lp - lm) / (2 * h)
    scale = np.abs(fd).max()
    assert scale > 1e-3
    assert np.abs(g - fd).max() < 1e-6 * scale + 1e-9


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(0)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    q, m, v = mlp.adam(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    assert np.allclose(q, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


@pytest.mark.parametrize("mol", [M.octatetraene(), M.batch_member(7)])
def test_features_direct_loop(mol):
    if mol.n > 24:
        mol = M.random_fragment(14, seed=2)
    p = prepare(mol)
    C, F, G, o = p["C"], p["F"], p["G"], p["nocc"]
    n = C.shape[0]
    v = n - o
    f, t, Dp = mlp.features(C, F, G, o)
    g = ccsd.eri(C, G)
    e = np.diag(F)
    mu = 0.5 * (e[o - 1] + e[o])
    rng = np.random.default_rng(1)
    for _ in range(40):
        i, j = rng.integers(0, o, 2)
        a, b = rng.integers(0, v, 2)
        A, B = o + a, o + b
        K = g[i, A, j, B]
        X = g[i, B, j, A]
        dp = e[A] + e[B] - e[i] - e[j]

        def ov(pq, rs):
            return n * sum(C[m, pq] ** 2 * C[m, rs] ** 2 for m in range(n))

        want = [10 * (-K / dp), K / 2, X / 2, dp / 10, 2 / dp, (e[i] + e[j] - 2 * mu) / 10,
                (e[A] + e[B] - 2 * mu) / 10, abs(e[i] - e[j]) / 10, abs(e[A] - e[B]) / 10,
                ov(i, j), ov(A, B), (ov(i, A) + ov(j, B)) / 2, abs(ov(i, A) - ov(j, B)) / 2,
                (ov(i, B) + ov(j, A)) / 2, abs(ov(i, B) - ov(j, A)) / 2, 10 / n]
        assert np.allclose(f[i, j, a, b], want, rtol=1e-10, atol=1e-12)
        assert np.isclose(t[i, j, a, b], -K / dp, rtol=1e-12)
    # pair symmetry: (i,a) <-> (j,b) leaves every feature unchanged
    assert np.allclose(f, f.transpose(1, 0, 3, 2, 4), rtol=1e-13, atol=1e-15)


def test_learned_guess_is_mp2_plus_correction():
    p = prepare(M.octatetraene())
    C, F, G, o = p["C"], p["F"], p["G"], p["nocc"]
    params = mlp.init(1)
    f, t, Dp = mlp.features(C, F, G, o)
    y, _ = mlp.forward(params, f.reshape(-1, mlp.NF).astype(np.float32))
    tl = mlp.guess(params, C, F, G, o)
    assert np.allclose(tl - t, y.reshape(t.shape) / Dp)
    z = np.zeros(mlp.NPARAM, dtype=np.float32)   # a zero network returns the MP2 amplitudes
    assert np.allclose(mlp.guess(z, C, F, G, o), t)
