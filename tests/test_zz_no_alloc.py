"""The allocation-free entry point (kmi_gemm_bf16_noalloc) returns the bits of kmi_gemm_bf16,
also on a second, smaller call that reuses the static workspace the first one left behind."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WORKSPACE = 32 * 1024 * 1024  # KMI_NOALLOC_WORKSPACE_BYTES


def _operands(M, N, K, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((M, K)).astype(np.float32), rng.standard_normal((K, N)).astype(np.float32)


def test_noalloc_gemm_matches_allocating_entry_point(kmi, backend):
    from kraken_mi355x.nn import gemm_bf16

    # first call: a split-K product; second call: smaller, with a transposed operand view
    for (M, N, K, splits, trans, seed) in [(300, 130, 1000, 4, False, 1), (70, 33, 96, 2, True, 2)]:
        print(f"[noalloc] M={M} N={N} K={K} splits={splits} trans={trans}", flush=True)
        A, B = _operands(M, N, K, seed)
        if trans:
            B = np.ascontiguousarray(B.T).T  # column-major view
        want = gemm_bf16(A, B, splits=splits, backend=backend)
        got = gemm_bf16(A, B, splits=splits, backend=backend, noalloc=True)
        assert np.array_equal(got, want)
        # and the product itself, to BF16 operand rounding
        ref = A.astype(np.float64) @ B.astype(np.float64)
        assert np.max(np.abs(got - ref)) <= 2e-2 * np.sqrt(K)


def test_noalloc_gemm_rejects_problems_beyond_the_workspace(kmi, backend):
    from kraken_mi355x import _lib as L

    M = N = K = 2048  # 6 (na + nb) + 4 M N bytes > 32 MiB
    assert 6 * (M * K + K * N) + 4 * M * N > WORKSPACE
    A = np.zeros((M, K), np.float32)
    B = np.zeros((K, N), np.float32)
    C = np.zeros((M, N), np.float32)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    rc = L.load().kmi_gemm_bf16_noalloc(backend.ctx, M, N, K, fp(A), K, 1, fp(B), N, 1, 1, fp(C))
    assert rc == L.KMI_ERR_ARG
    assert b"workspace" in L.load().kmi_last_error(backend.ctx)
