"""Learned T2 guess on the GPU against the oracle (oracle/mlp.py).

Tolerances (BF16 matrix-core products, FP32 accumulation and erff on the
device; float64 with the same BF16 roundings on the CPU; a GELU value that
sits on a BF16 rounding tie may round the other way):
  network output     |dy| <= 2e-3 max|y| + 1e-4
  gradient           max|dg| <= 1e-2 max|g|, loss to 1e-4 relative
  features           1e-5 relative (FP64 on both sides, rounded to FP32)
  learned amplitudes |dt2| <= 1e-3 max|t2 - t_mp2| + 1e-7, pair symmetry exact
"""
import numpy as np
import pytest

from helios_amd import capi, molecules as M
from oracle import mlp
from tests._molprep import prepare

pytestmark = pytest.mark.gpu


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = mlp.init(seed).astype(np.float64)
    p[mlp.OFF_W3:mlp.OFF_B3] *= 20.0
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.normal(0, 0.3, 64)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.normal(0, 0.3, 64)
    p[mlp.OFF_B3] = 0.1
    return p.astype(np.float32)


def test_forward_matches_oracle():
    rng = np.random.default_rng(0)
    p = random_params(1)
    X = rng.normal(0, 1, (1000, mlp.NF)).astype(np.float32)
    tr = capi.Trainer(16)
    tr.set_params(p)
    y = tr.predict(X)
    yo, _ = mlp.forward(p, X)
    assert np.abs(y - yo).max() <= 2e-3 * np.abs(yo).max() + 1e-4


@pytest.mark.parametrize("count", [100, 700])
def test_gradient_matches_oracle(count):
    rng = np.random.default_rng(count)
    p = random_params(2)
    rows = np.zeros((count, mlp.NF + 2), dtype=np.float32)
    rows[:, :mlp.NF] = rng.normal(0, 1, (count, mlp.NF))
    rows[:, mlp.NF] = rng.normal(0, 0.5, count)
    rows[:, mlp.NF + 1] = rng.uniform(0.0, 1.0, count)
    tr = capi.Trainer(count)
    tr.set_params(p)
    tr.write_samples(rows)
    loss, g = tr.gradient()
    lo, go = mlp.backward(p, rows[:, :mlp.NF], rows[:, mlp.NF].astype(np.float64),
                          rows[:, mlp.NF + 1].astype(np.float64))
    assert abs(loss - lo) <= 1e-4 * lo
    assert np.abs(g - go).max() <= 1e-2 * np.abs(go).max()
    # per-block checks so a wrong layout in a small block cannot hide behind a large one
    for a, b in [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
                 (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.NPARAM)]:
        assert np.abs(g[a:b] - go[a:b]).max() <= 1e-2 * np.abs(go[a:b]).max() + 1e-7


def test_adam_steps_reproduce_and_follow_the_gradient():
    rng = np.random.default_rng(3)
    p = random_params(3)
    n = 512
    rows = np.zeros((n, mlp.NF + 2), dtype=np.float32)
    rows[:, :mlp.NF] = rng.normal(0, 1, (n, mlp.NF))
    rows[:, mlp.NF] = rng.normal(0, 0.5, n)
    rows[:, mlp.NF + 1] = 1.0
    runs = []
    for _ in range(2):
        tr = capi.Trainer(n)
        tr.set_params(p)
        tr.write_samples(rows)
        hist = tr.train(epochs=4, batch=128, lr=1e-3, lr_final=1e-4)
        runs.append((hist, tr.params()))
        tr.close()
    # no atomics: bitwise reproducible
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][0][-1] < runs[0][0][0]
    # first Adam step from zero moments: w1 - w0 = -lr g / (|g| + 1e-8), i.e. -lr sign(g)
    tr = capi.Trainer(n)
    tr.set_params(p)
    tr.write_samples(rows)
    _, g = tr.gradient()
    tr.train(epochs=1, batch=n, lr=1e-3, lr_final=1e-3)
    step = tr.params().astype(np.float64) - p.astype(np.float64)
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    assert big.sum() > 0.9 * mlp.NPARAM
    want = -1e-3 * np.sign(g[big])
    assert np.abs(step[big] - want).max() < 1e-3 * 1e-3 + 1e-6 * np.abs(p[big]).max()
    assert np.abs(step).max() <= 1e-3 * (1 + 1e-5) + 1e-6 * np.abs(p).max()


def solved_context(mol):
    p = prepare(mol)
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    ctx.set_molecule(p["C"], p["F"], p["G"], p["nocc"], p["eref"])
    e, it, st = ctx.solve(capi.GUESS_MP2, 300, 1e-10)
    assert st == 1
    return p, ctx


@pytest.mark.parametrize("mol", [M.octatetraene(), M.random_fragment(16, seed=4)])
def test_collected_features_and_targets(mol):
    p, ctx = solved_context(mol)
    o, v = p["nocc"], p["n"] - p["nocc"]
    tr = capi.Trainer(300)
    assert tr.collect(ctx, 300, seed=7) == 300
    rows, where = tr.read_samples()
    assert (where[:, 0] == 0).all() and (where[:, 1] >= 0).all() and (where[:, 1] < o * o * v * v).all()
    f, t, Dp = mlp.features(p["C"], p["F"], p["G"], o)
    _, t2 = ctx.amplitudes()
    for r, e in zip(rows, where[:, 1]):
        i, j, a, b = np.unravel_index(e, (o, o, v, v))
        assert np.allclose(r[:mlp.NF], f[i, j, a, b], rtol=1e-5, atol=1e-7)
        want = Dp[i, j, a, b] * (t2[i, j, a, b] - t[i, j, a, b])
        assert abs(r[mlp.NF] - want) <= 1e-5 * abs(want) + 1e-9
        assert r[mlp.NF + 1] == 1.0


@pytest.mark.parametrize("mol", [M.octatetraene(), M.random_fragment(16, seed=4)])
def test_learned_guess_kernel(mol):
    p = prepare(mol)
    o = p["nocc"]
    params = random_params(5)
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    ctx.set_molecule(p["C"], p["F"], p["G"], o, p["eref"])
    with pytest.raises(capi.HccError):
        ctx.guess(capi.GUESS_LEARNED)        # no network yet: a clear error, no silent fallback
    ctx.set_guess_network(params)
    ctx.guess(capi.GUESS_LEARNED)
    _, t2 = ctx.amplitudes()
    want = mlp.guess(params, p["C"], p["F"], p["G"], o)
    _, tm, _ = mlp.features(p["C"], p["F"], p["G"], o)
    assert np.array_equal(t2, t2.transpose(1, 0, 3, 2))
    assert np.abs(t2 - want).max() <= 1e-3 * np.abs(want - tm).max() + 1e-7
    # and the solve from it reaches the same energy as from MP2
    e1, it1, st1 = ctx.solve(capi.GUESS_LEARNED, 300, 1e-10)
    e2, it2, st2 = ctx.solve(capi.GUESS_MP2, 300, 1e-10)
    assert st1 == 1 and st2 == 1 and abs(e1[0] - e2[0]) < 1e-8


def test_training_on_solver_amplitudes_lowers_loss():
    mols = [prepare(M.batch_member(k, base_seed=2_000_000)) for k in range(24)]
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    ctx.batch_load([dict(C=m["C"], F=m["F"], G=m["G"], nocc=m["nocc"], eref=m["eref"]) for m in mols])
    ctx.batch_solve(capi.GUESS_MP2, 300, 1e-9)
    tr = capi.Trainer(24 * 128)
    tr.set_params(mlp.init(0))
    tr.collect(ctx, 128, seed=1)
    l0 = tr.evaluate()
    hist = tr.train(epochs=20, batch=512, lr=3e-3, lr_final=1e-4)
    l1 = tr.evaluate()
    assert l1 < 0.7 * l0 and hist[-1] < hist[0]
