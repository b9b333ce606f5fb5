"""GPU parity of the two-phase SpGEMM (bclx_spgemm_symbolic + bclx_spgemm_numeric) against the oracle.

Structure (row_ptr, col_idx) must be bit-identical to the oracle's restatement of BCL::gemm(CSR, CSR, CSR)
(bcl/containers/algorithms/spgemm.hpp:40-76), which keeps exact-zero sums. Values are accumulated in float64
and rounded once, so each entry must be within one rounding of the float64 oracle:
|c - c64| <= 2^-24 |c64| + 1e-13 sum|a||b| for fp32 (2^-52 for fp64), far inside the 1e-5 relative target.
"""
import numpy as np
import pytest
import torch

import oracle_py as O
from bcl_amd import gen
from bcl_amd.csr import CSR
from bcl_amd.device import DeviceCSR, spgemm
from conftest import spgemm_value_bound

pytestmark = pytest.mark.gpu


def _abs(a: CSR) -> CSR:
    return CSR(a.m, a.n, a.row_ptr, a.col_idx, np.abs(a.vals))


def check_product(handle, a: CSR, b: CSR, ptr_dtype=torch.int32, out_ptr=torch.int64):
    print(f"[spgemm] A {a.m}x{a.n} nnz={a.nnz} B {b.m}x{b.n} nnz={b.nnz} {a.vals.dtype}", flush=True)
    A = DeviceCSR.from_host(a, ptr_dtype=ptr_dtype)
    B = DeviceCSR.from_host(b, ptr_dtype=ptr_dtype)
    C = spgemm(handle, A, B, ptr_dtype=out_ptr)
    torch.cuda.synchronize()
    c = C.to_host()
    ref = O.spgemm(a, b)
    assert c.nnz == ref.nnz
    np.testing.assert_array_equal(c.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(c.col_idx, ref.col_idx)
    c64 = O.spgemm_f64(a, b)
    absum = O.spgemm_f64(_abs(a), _abs(b))
    np.testing.assert_array_equal(c64.col_idx, ref.col_idx)
    err = np.abs(c.vals.astype(np.float64) - c64.vals)
    bound = spgemm_value_bound(a, b, c64.vals, absum.vals, a.vals.dtype)
    bad = np.nonzero(err > bound)[0]
    assert bad.size == 0, f"{bad.size} values outside the bound, first at {bad[:5]}: {c.vals[bad[:5]]} vs {c64.vals[bad[:5]]}"
    return c, ref


def test_uniform_small(handle):
    a = gen.uniform(2048, 2048, 8, seed=1)
    check_product(handle, a, a)


def test_uniform_config1(handle):
    """16384 x 16384, exactly 8 nnz/row, C = A * A (the plain single-tile workload)."""
    a = gen.uniform(16384, 16384, 8, seed=0xB0C1)
    c, _ = check_product(handle, a, a)
    assert c.nnz > 1_000_000


def test_rectangular_and_empty_rows(handle):
    a = gen.uniform(300, 1000, 5, seed=3)
    b = gen.banded(1000, 700, bw=40, max_per_row=9, seed=4)
    # blank out some rows of a and b
    keep_a = np.ones(a.m, bool)
    keep_a[::7] = False
    rows = np.repeat(np.arange(a.m), a.row_nnz())
    mask = keep_a[rows]
    rp = np.zeros(a.m + 1, np.int64)
    rp[1:] = np.cumsum(np.bincount(rows[mask], minlength=a.m))
    a2 = CSR(a.m, a.n, rp, a.col_idx[mask], a.vals[mask])
    check_product(handle, a2, b)


def test_rmat_heavy_rows(handle):
    """R-MAT scale 14: hub rows exceed the LDS tables and take the global accumulator."""
    a = gen.rmat(14, seed=5)
    _, prod = O.row_nnz(a, a)
    assert (np.minimum(prod, a.n) > 4096).any(), "test needs heavy rows"
    check_product(handle, a, a)


def test_rmat_sliced_heavy_rows(handle):
    """R-MAT scale 15 (32768 columns): hub rows have more output entries than one LDS slice of the heavy-row kernel
    (several slices, each restricted to its column range) and more A entries than one staged chunk."""
    a = gen.rmat(15, seed=7)
    nnz, _ = O.row_nnz(a, a)
    assert (nnz > 9216).sum() > 100, "test needs rows spanning several slices"
    assert (np.diff(a.row_ptr) > 1024).any(), "test needs rows with more than one chunk of A entries"
    check_product(handle, a, a)


def test_zero_retention(handle):
    """A = [[1, 1]], B = [[1], [-1]]: C keeps the exact-zero entry (spgemm.hpp:69 never erases)."""
    a = CSR(1, 2, np.array([0, 2]), np.array([0, 1], np.int32), np.array([1, 1], np.float32))
    b = CSR(2, 1, np.array([0, 1, 2]), np.array([0, 0], np.int32), np.array([1, -1], np.float32))
    c, ref = check_product(handle, a, b)
    assert c.nnz == 1 and c.vals[0] == 0.0


def test_empty_operands(handle):
    a = CSR.empty(64, 32)
    b = gen.uniform(32, 48, 3, seed=2)
    c, _ = check_product(handle, a, b)
    assert c.nnz == 0
    c, _ = check_product(handle, b.astype(np.float32), CSR.empty(48, 16))
    assert c.nnz == 0


def test_int64_row_ptr(handle):
    a = gen.rmat(10, seed=9)
    check_product(handle, a, a, ptr_dtype=torch.int64, out_ptr=torch.int32)


def test_laplacian_fp64(handle):
    """2D 5-point Laplacian squared in float64: integer values are exact."""
    a = gen.laplacian2d(40)
    c, ref = check_product(handle, a, a)
    np.testing.assert_array_equal(c.vals, ref.vals)


def test_deterministic_structure(handle):
    a = gen.rmat(12, seed=11)
    A = DeviceCSR.from_host(a)
    c1 = spgemm(handle, A, A).to_host()
    c2 = spgemm(handle, A, A).to_host()
    np.testing.assert_array_equal(c1.row_ptr, c2.row_ptr)
    np.testing.assert_array_equal(c1.col_idx, c2.col_idx)
