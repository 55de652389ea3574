"""Grouped GEMM engine against numpy (FP64 MFMA and BF16 MFMA paths)."""
import numpy as np
import pytest

from helios_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 4), (64, 64, 16), (65, 63, 17), (100, 37, 250),
                                   (5, 300, 7), (257, 129, 64)])
def test_gemm_f64(m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((k, n))
    C0 = rng.standard_normal((m, n))
    C = capi.selftest_gemm(A, B, C0, alpha=0.75, beta=-0.5)
    ref = 0.75 * A @ B - 0.5 * C0
    err = float(np.abs(C - ref).max())
    assert err < 1e-12 * max(1.0, float(np.abs(ref).max())) * max(1, k) ** 0.5, err


def test_gemm_f64_beta_zero_ignores_garbage():
    A = np.ones((20, 8))
    B = np.ones((8, 30))
    C0 = np.full((20, 30), np.nan)
    C = capi.selftest_gemm(A, B, C0, alpha=1.0, beta=0.0)
    assert np.all(C == 8.0)


def _bf16(x):
    """Round to BF16 through float32 (round to nearest even), as the kernel does."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(a.shape)


@pytest.mark.parametrize("tile", [32, 64, 128])
@pytest.mark.parametrize("a_kfast", [True, False])
@pytest.mark.parametrize("b_nfast", [True, False])
@pytest.mark.parametrize("m,n,k,which", [(64, 64, 32, "AB"), (70, 50, 100, "AB"), (5, 300, 7, "AB"),
                                         (129, 65, 600, "AB"), (70, 50, 100, "A"), (70, 50, 100, "B"),
                                         (300, 257, 45, "AB")])
def test_gemm_correction(m, n, k, which, a_kfast, b_nfast, tile):
    """Cd += alpha (A0 dB + dA (B0 + dB)): operands rounded to BF16 (B0 + dB
    summed in FP32 first), FP32 accumulation, against float64 products of the
    same rounded operands."""
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A0 = rng.standard_normal((m, k))
    B0 = rng.standard_normal((k, n))
    dA = (1e-2 * rng.standard_normal((m, k))).astype(np.float32) if "A" in which else None
    dB = (1e-2 * rng.standard_normal((k, n))).astype(np.float32) if "B" in which else None
    C0 = rng.standard_normal((m, n)).astype(np.float32)
    C, _ = capi.selftest_gemm_correction(A0, dA, B0, dB, C0, alpha=0.5, beta=-1.0, a_kfast=a_kfast, b_nfast=b_nfast,
                                         tile=tile)
    b0 = _bf16(B0)
    terms, mags = [], []
    if dB is not None:
        a0, db = _bf16(A0), _bf16(dB)
        terms.append(a0 @ db)
        mags.append(np.abs(a0) @ np.abs(db))
    if dA is not None:
        da = _bf16(dA)
        b1 = _bf16((b0.astype(np.float32) + (dB if dB is not None else 0.0)).astype(np.float32))
        terms.append(da @ b1)
        mags.append(np.abs(da) @ np.abs(b1))
    ref = 0.5 * sum(terms) - C0.astype(np.float64)
    bound = 0.5 * sum(mags) * 2e-6 * np.sqrt(k) + np.abs(ref) * 2 ** -23
    assert np.all(np.abs(C - ref) <= bound + 1e-30), float(np.max(np.abs(C - ref) / (bound + 1e-30)))
