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


@pytest.mark.parametrize("m,n,k", [(64, 64, 32), (70, 50, 100)])
def test_gemm_bf16(m, n, k):
    rng = np.random.default_rng(k)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((k, n))
    C = capi.selftest_gemm(A, B, precision=capi.HCC_MIXED)
    ref = A @ B
    # BF16 inputs (8-bit significand), FP32 accumulation
    err = float(np.abs(C - ref).max())
    assert err < 2e-2 * np.sqrt(k), err
