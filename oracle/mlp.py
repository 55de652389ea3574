"""CPU restatement of the learned T2 guess (test infrastructure only).

The network has no counterpart in the reference: it replaces the guess of
src/methods/coupledCluster/mod_cc.f90:634-650 by the MP2 amplitudes plus a
learned correction.  Everything here restates the device code of
helios_amd/csrc/hcc_mlp.hip in float64 numpy, with the same BF16 roundings
where the device rounds, so that the tests can compare the two:

  features(...)   the 16 per-amplitude inputs
  forward(...)    3-layer MLP (16 -> 64 -> 64 -> 1, exact-erf GELU);
                  weights W1, W2, W3 and the layer inputs x, h1 are rounded
                  to BF16 before the products (the matrix cores multiply
                  BF16), h2 and the biases stay FP32, sums are exact here
  backward(...)   the analytic gradient of the mean squared error, with the
                  roundings treated as the identity (straight-through)
  adam(...)       the optimiser step

Pinning (tests/test_mlp.py): the gradient of the unrounded network is
checked against central finite differences in float64, element by element;
the features against a direct loop over the MO integrals of oracle.ccsd.eri
and against exact pair symmetry; bf16() against hand-computed bit patterns
(ties to even, overflow-free range).
"""
import numpy as np

NF = 16          # features
NH = 64          # hidden width
# flat parameter layout (float32): W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3
OFF_W1, OFF_B1 = 0, 1024
OFF_W2, OFF_B2 = 1088, 5184
OFF_W3, OFF_B3 = 5248, 5312
NPARAM = 5313


def bf16(x):
    """Round float32 values to BF16 (round to nearest, ties to even); returns float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(NH, NF), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(NH, NH), b2=p[OFF_B2:OFF_W3],
                W3=p[OFF_W3:OFF_B3], b3=p[OFF_B3:NPARAM])


def pack(d):
    p = np.zeros(NPARAM)
    p[OFF_W1:OFF_B1] = d["W1"].ravel()
    p[OFF_B1:OFF_W2] = d["b1"]
    p[OFF_W2:OFF_B2] = d["W2"].ravel()
    p[OFF_B2:OFF_W3] = d["b2"]
    p[OFF_W3:OFF_B3] = d["W3"]
    p[OFF_B3:NPARAM] = d["b3"]
    return p


def init(seed=0):
    """He-style initial weights (the device trainer starts from these)."""
    rng = np.random.default_rng(seed)
    d = dict(W1=rng.normal(0, np.sqrt(2.0 / NF), (NH, NF)), b1=np.zeros(NH),
             W2=rng.normal(0, np.sqrt(2.0 / NH), (NH, NH)), b2=np.zeros(NH),
             W3=rng.normal(0, np.sqrt(1.0 / NH), NH) * 0.1, b3=np.zeros(1))
    return pack(d).astype(np.float32)


def _erf(x):
    from scipy.special import erf
    return erf(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / np.sqrt(2.0)))


def gelu_prime(z):
    return 0.5 * (1.0 + _erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, X, rounding=True):
    """X: (n, 16) features.  Returns y (n,) and the cache for backward()."""
    w = unpack(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, W2, W3 = r(w["W1"]), r(w["W2"]), r(w["W3"])
    x = r(X)
    z1 = x @ W1.T + w["b1"]
    h1 = gelu(z1)
    h1r = r(h1)
    z2 = h1r @ W2.T + w["b2"]
    h2 = gelu(z2)
    y = h2 @ W3 + w["b3"][0]
    return y, dict(x=x, z1=z1, h1r=h1r, z2=z2, h2=h2, W2=W2, W3=W3)


def loss(y, target, weight=None):
    """Mean over the rows of weight * (y - target)^2."""
    wt = np.ones_like(y) if weight is None else weight
    return float(np.mean(wt * (y - target) ** 2))


def backward(p, X, target, weight=None, rounding=True):
    """Loss and gradient (flat, NPARAM) of loss(forward(p, X), target)."""
    y, c = forward(p, X, rounding)
    n = len(y)
    wt = np.ones(n) if weight is None else weight
    dy = 2.0 * wt * (y - target) / n
    dz2 = (dy[:, None] * c["W3"][None, :]) * gelu_prime(c["z2"])
    dh1 = dz2 @ c["W2"]
    dz1 = dh1 * gelu_prime(c["z1"])
    g = dict(W1=dz1.T @ c["x"], b1=dz1.sum(0), W2=dz2.T @ c["h1r"], b2=dz2.sum(0),
             W3=c["h2"].T @ dy, b3=np.array([dy.sum()]))
    return loss(y, target, wt), pack(g)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); returns the new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ----------------------------------------------------------------------------
# features
# ----------------------------------------------------------------------------
def overlaps(C):
    """O[p][q] = n * sum_mu C_mu,p^2 C_mu,q^2 (site-localisation overlap of two MOs)."""
    n = C.shape[0]
    S = C * C
    return n * (S.T @ S)


def features(C, F, G, o):
    """Features (o, o, v, v, 16), MP2 amplitudes (o,o,v,v) and Dp = e_a+e_b-e_i-e_j.

    With e_p = F[p,p], mu = (e_HOMO + e_LUMO)/2, K = (ia|jb), X = (ib|ja):
      0  10 t_mp2 (t_mp2 = K / (e_i + e_j - e_a - e_b))
      1  K / 2              2  X / 2
      3  Dp / 10            4  2 / Dp
      5  (e_i + e_j - 2 mu) / 10     6  (e_a + e_b - 2 mu) / 10
      7  |e_i - e_j| / 10            8  |e_a - e_b| / 10
      9  O_ij    10  O_ab    11  (O_ia + O_jb) / 2    12  |O_ia - O_jb| / 2
      13 (O_ib + O_ja) / 2   14  |O_ib - O_ja| / 2    15  10 / n
    Every feature is invariant under (i,a) <-> (j,b).
    """
    from oracle.ccsd import eri
    n = C.shape[0]
    v = n - o
    e = np.diag(F)
    mu = 0.5 * (e[o - 1] + e[o])
    g = eri(C, G)
    K = g[:o, o:, :o, o:].transpose(0, 2, 1, 3)          # (i a | j b) -> [i, j, a, b]
    X = g[:o, o:, :o, o:].transpose(0, 2, 3, 1)          # (i b | j a) -> [i, j, a, b]
    ei = e[:o][:, None, None, None]
    ej = e[:o][None, :, None, None]
    ea = e[o:][None, None, :, None]
    eb = e[o:][None, None, None, :]
    Dp = ea + eb - ei - ej
    t = -K / Dp
    O = overlaps(C)
    Oij = O[:o, :o][:, :, None, None]
    Oab = O[o:, o:][None, None, :, :]
    Oia = O[:o, o:][:, None, :, None]
    Ojb = O[:o, o:][None, :, None, :]
    Oib = O[:o, o:][:, None, None, :]
    Oja = O[:o, o:][None, :, :, None]
    shape = (o, o, v, v)
    f = np.empty(shape + (NF,))
    f[..., 0] = 10.0 * t
    f[..., 1] = K / 2
    f[..., 2] = X / 2
    f[..., 3] = Dp / 10
    f[..., 4] = 2.0 / Dp
    f[..., 5] = np.broadcast_to((ei + ej - 2 * mu) / 10, shape)
    f[..., 6] = np.broadcast_to((ea + eb - 2 * mu) / 10, shape)
    f[..., 7] = np.broadcast_to(np.abs(ei - ej) / 10, shape)
    f[..., 8] = np.broadcast_to(np.abs(ea - eb) / 10, shape)
    f[..., 9] = np.broadcast_to(Oij, shape)
    f[..., 10] = np.broadcast_to(Oab, shape)
    f[..., 11] = (Oia + Ojb) / 2
    f[..., 12] = np.abs(Oia - Ojb) / 2
    f[..., 13] = (Oib + Oja) / 2
    f[..., 14] = np.abs(Oib - Oja) / 2
    f[..., 15] = 10.0 / n
    return f, t, Dp


def guess(p, C, F, G, o):
    """Learned T2 guess: t_mp2 + y / Dp with y the network output (float64 here)."""
    f, t, Dp = features(C, F, G, o)
    y, _ = forward(p, f.reshape(-1, NF).astype(np.float32))
    return t + y.reshape(t.shape) / Dp
