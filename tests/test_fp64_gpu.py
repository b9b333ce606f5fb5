"""The fp64 overloads at the boundary: dense GEMM on the fp64 matrix cores (cblas_gemm_wrapper_'s double overload ->
cblas_dgemm, bcl/containers/algorithms/cblas_wrapper.hpp:29-40) and SpMM in fp64 (the reference's spmm_wrapper on
double, bcl/containers/algorithms/spgemm.hpp:119-149), against float64 references on the CPU (numpy's matrix product,
scipy's sparse product).

Bounds: a float64 result summed in any order is within gamma_K * sum_k |a_ik||b_kj| of the exact product
(gamma_K = K u / (1 - K u), u = 2^-53), so |GPU - CPU| <= 2 gamma_K sum|a||b| when both are float64 sums; values that
are small integers are exact and must match bit for bit.
"""
import numpy as np
import pytest
import torch

from bcl_amd import gen
from bcl_amd.csr import CSR
from bcl_amd.device import DeviceCSR, gemm_f64, spmm

pytestmark = pytest.mark.gpu
U64 = 2.0 ** -53


def sp(a: CSR):
    """a as a scipy CSR matrix (float64 sums in scipy's own order: the float64 reference of the SpMM tests)."""
    import scipy.sparse
    return scipy.sparse.csr_matrix((a.vals, a.col_idx, a.row_ptr), shape=(a.m, a.n))


def gamma(k):
    return k * U64 / (1 - k * U64)


@pytest.mark.parametrize("m,n,k,ta,tb", [(257, 190, 333, False, False), (128, 128, 16, True, False),
                                        (300, 70, 1000, False, True), (64, 1, 77, True, True), (1024, 1024, 1024, False, False)])
def test_gemm_f64_random(handle, m, n, k, ta, tb):
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    opA, opB = (A.T if ta else A), (B.T if tb else B)
    alpha, beta = 0.75, -1.5
    want = beta * C0 + alpha * (opA @ opB)
    S = np.abs(beta * C0) + abs(alpha) * (np.abs(opA) @ np.abs(opB))
    dC = torch.from_numpy(C0).cuda()
    gemm_f64(handle, torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, trans_a=ta, trans_b=tb, alpha=alpha,
             beta=beta)
    got = dC.cpu().numpy()
    err = np.abs(got - want)
    assert (err <= 2 * gamma(k + 2) * S + 1e-300).all(), float((err / S).max())


def test_gemm_f64_integers_exact(handle):
    """fill_range-like small integers (DMatrix.hpp:589-604): every partial sum is an exact integer in fp64."""
    m = n = k = 384
    A = (np.arange(m * k) % 13 - 6).reshape(m, k).astype(np.float64)
    B = (np.arange(k * n) % 7 - 3).reshape(k, n).astype(np.float64)
    dC = torch.ones((m, n), dtype=torch.float64, device="cuda")
    gemm_f64(handle, torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC)  # C += A B (alpha = beta = 1)
    np.testing.assert_array_equal(dC.cpu().numpy(), 1.0 + A @ B)


def test_gemm_f64_8192_rows(handle):
    """8192^3 in fp64: 16 rows of C against numpy's float64 product of the same rows."""
    N = 8192
    g = torch.Generator(device="cuda").manual_seed(21)
    A = torch.randn((N, N), device="cuda", dtype=torch.float64, generator=g)
    B = torch.randn((N, N), device="cuda", dtype=torch.float64, generator=g)
    C = torch.zeros((N, N), device="cuda", dtype=torch.float64)
    gemm_f64(handle, A, B, C, beta=0.0)
    torch.cuda.synchronize()
    rows = np.arange(0, N, N // 16)
    Ah, Bh = A[torch.from_numpy(rows).cuda()].cpu().numpy(), B.cpu().numpy()
    want = Ah @ Bh
    S = np.abs(Ah) @ np.abs(Bh)
    err = np.abs(C[torch.from_numpy(rows).cuda()].cpu().numpy() - want)
    assert (err <= 2 * gamma(N) * S).all()
    print(f"[gemm64] 8192^3: max error / sum|a||b| {float((err / S).max()):.3e}", flush=True)


@pytest.mark.parametrize("ptr64", [False, True])
def test_spmm_f64(handle, ptr64):
    a = gen.rmat(14, seed=3)
    a = CSR(a.m, a.n, a.row_ptr, a.col_idx, a.vals.astype(np.float64))
    n, ld = 96, 100
    rng = np.random.default_rng(4)
    B = rng.standard_normal((a.n, ld))
    C0 = rng.standard_normal((a.m, ld))
    A = DeviceCSR.from_host(a, ptr_dtype=torch.int64 if ptr64 else torch.int32)
    dC = torch.from_numpy(C0).cuda()
    spmm(handle, A, torch.from_numpy(B).cuda(), dC, n=n, alpha=1.0, beta=1.0)
    got = dC.cpu().numpy()
    want = C0[:, :n] + sp(a) @ B[:, :n]
    S = np.abs(C0[:, :n]) + sp(CSR(a.m, a.n, a.row_ptr, a.col_idx, np.abs(a.vals))) @ np.abs(B[:, :n])
    L = np.diff(a.row_ptr)[:, None] + 2
    assert (np.abs(got[:, :n] - want) <= 2 * gamma(L) * S + 1e-300).all()
    np.testing.assert_array_equal(got[:, n:], C0[:, n:])


def test_spmm_f64_integers_exact(handle):
    a = gen.uniform(3000, 2000, 9, seed=5)
    a = CSR(a.m, a.n, a.row_ptr, a.col_idx, np.round(a.vals * 4).astype(np.float64))
    B = (np.arange(a.n * 64) % 11 - 5).reshape(a.n, 64).astype(np.float64)
    dC = torch.zeros((a.m, 64), dtype=torch.float64, device="cuda")
    spmm(handle, DeviceCSR.from_host(a), torch.from_numpy(B).cuda(), dC)
    np.testing.assert_array_equal(dC.cpu().numpy(), a.to_dense() @ B)
