"""GPU parity of the dense tile GEMM (bclx_gemm; cblas_gemm_wrapper_ semantics, row-major) against the oracle.

Error bounds against the float64 oracle, per entry, with S = sum_k |a_ik b_kj|:
  FP32 MFMA        |c - c64| <= 2e-6 * S      (fp32 products exact, fp32 accumulation, K <= 1024 here)
  BF16 MFMA        |c - c64| <= 2^-7 * S      (each input rounded to 8 significant bits)
  split-BF16 (3)   |c - c64| <= 2e-5 * S      (dropped lo*lo term and lo rounding ~2^-16, plus fp32 accumulation)
fill_range inputs (bcl/containers/DMatrix.hpp:589-604, small integers) are exact in every mode.
The large BF16 cases run the 256 x 128 block tiling (the one the 8192^3 product uses) with every transpose, ragged
edges and a padded leading dimension.
"""
import numpy as np
import pytest
import torch

import oracle_py as O
from bcl_amd.device import GEMM_BF16, GEMM_FP32, GEMM_SPLIT_BF16, gemm, gemm_f64_ref

pytestmark = pytest.mark.gpu
BOUND = {GEMM_FP32: 2e-6, GEMM_BF16: 2.0 ** -7, GEMM_SPLIT_BF16: 2e-5}


def run(handle, a, b, c0, mode, ta=False, tb=False, alpha=1.0, beta=1.0):
    A, B, C = (torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (a, b, c0))
    gemm(handle, A, B, C, mode=mode, trans_a=ta, trans_b=tb, alpha=alpha, beta=beta)
    torch.cuda.synchronize()
    return C.cpu().numpy()


@pytest.mark.parametrize("mode", [GEMM_FP32, GEMM_BF16, GEMM_SPLIT_BF16])
def test_fill_range_exact(handle, mode):
    M, N, K = 256, 384, 192
    a = (np.arange(M * K) % 71).reshape(M, K).astype(np.float32)
    b = (np.arange(K * N) % 121).reshape(K, N).astype(np.float32)
    c = run(handle, a, b, np.zeros((M, N), np.float32), mode)
    np.testing.assert_array_equal(c, (a.astype(np.int64) @ b.astype(np.int64)).astype(np.float32))


@pytest.mark.parametrize("mode", [GEMM_FP32, GEMM_BF16, GEMM_SPLIT_BF16])
@pytest.mark.parametrize("shape,ta,tb", [((512, 512, 512), False, False), ((300, 200, 700), True, False),
                                         ((129, 257, 65), False, True), ((64, 96, 1024), True, True)])
def test_random_vs_f64(handle, mode, shape, ta, tb):
    M, N, K = shape
    rng = np.random.default_rng(M + N + K)
    a = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    b = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    c0 = rng.standard_normal((M, N)).astype(np.float32)
    alpha, beta = 1.5, -0.5
    c = run(handle, a, b, c0, mode, ta, tb, alpha, beta)
    oa = np.ascontiguousarray(a.T if ta else a)
    ob = np.ascontiguousarray(b.T if tb else b)
    ref = beta * c0.astype(np.float64) + alpha * O.gemm_f64(oa, ob)
    S = abs(beta) * np.abs(c0) + abs(alpha) * O.gemm_f64(np.abs(oa), np.abs(ob))
    err = np.abs(c - ref)
    worst = float((err / S).max())
    print(f"[gemm] mode {mode} {shape} ta={ta} tb={tb}: max |c-c64|/S = {worst:.3e}")
    assert worst <= BOUND[mode]


def test_split_is_more_accurate_than_bf16(handle):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((256, 512)).astype(np.float32)
    b = rng.standard_normal((512, 256)).astype(np.float32)
    z = np.zeros((256, 256), np.float32)
    ref = O.gemm_f64(a, b)
    e = {m: np.abs(run(handle, a, b, z, m) - ref).max() for m in (GEMM_FP32, GEMM_BF16, GEMM_SPLIT_BF16)}
    assert e[GEMM_SPLIT_BF16] < e[GEMM_BF16] / 50
    assert e[GEMM_FP32] <= e[GEMM_SPLIT_BF16] * 4


def test_beta_zero_ignores_nan_and_k_zero(handle):
    a = np.ones((128, 0), np.float32)
    b = np.ones((0, 128), np.float32)
    c0 = np.full((128, 128), 3.0, np.float32)
    np.testing.assert_array_equal(run(handle, a, b, c0, GEMM_FP32, beta=2.0), 6.0)
    a = np.ones((130, 40), np.float32)
    b = np.ones((40, 70), np.float32)
    c0 = np.full((130, 70), np.nan, np.float32)
    np.testing.assert_array_equal(run(handle, a, b, c0, GEMM_BF16, beta=0.0), 40.0)


def test_f64_reference_kernel(handle):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((300, 200)).astype(np.float32)
    b = rng.standard_normal((200, 250)).astype(np.float32)
    c = gemm_f64_ref(handle, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()).cpu().numpy()
    np.testing.assert_allclose(c, O.gemm_f64(a, b), rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("ta,tb,pad", [(False, False, 0), (True, False, 0), (False, True, 0), (True, True, 0),
                                       (False, False, 3)])
def test_bf16_large(handle, ta, tb, pad):
    """4100 x 4130 x 1000 (17 x 17 blocks of 256, ragged in M, N and K), alpha / beta, optional ld padding."""
    M, N, K = 4100, 4130, 1000
    g = torch.Generator(device="cuda").manual_seed(M + 2 * ta + tb + pad)
    rows_a, cols_a = (K, M) if ta else (M, K)
    rows_b, cols_b = (N, K) if tb else (K, N)
    A = torch.randn((rows_a, cols_a + pad), device="cuda", generator=g)[:, :cols_a]
    B = torch.randn((rows_b, cols_b + pad), device="cuda", generator=g)[:, :cols_b]
    C0 = torch.randn((M, N), device="cuda", generator=g)
    C = C0.clone()
    gemm(handle, A, B, C, mode=GEMM_BF16, trans_a=ta, trans_b=tb, alpha=1.25, beta=-0.5)
    oa = (A.t() if ta else A).contiguous()
    ob = (B.t() if tb else B).contiguous()
    ref = gemm_f64_ref(handle, oa, ob) * 1.25 - 0.5 * C0.double()
    S = gemm_f64_ref(handle, oa.abs(), ob.abs()) * 1.25 + 0.5 * C0.double().abs()
    worst = float(((C.double() - ref).abs() / S).max())
    print(f"[gemm] bf16 4100 x 4130 x 1000 ta={ta} tb={tb} pad={pad}: max |c-c64|/S = {worst:.3e}", flush=True)
    assert worst <= BOUND[GEMM_BF16]
