"""The C++ host layer (include/bclx/bclx.hpp) driven like BCL code: CSRMatrix::dot, gemm (c += a * b),
SparseHashAccumulator, spmm_wrapper and cblas_gemm_wrapper_ (row- and column-major), through files in BCL's binary
CSR format. Results are checked against the oracle with the same bounds as the C-ABI tests."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
from bcl_amd import gen
from bcl_amd.csr import CSR, read_binary, write_binary

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "host_layer_driver")


def _abs(a):
    return CSR(a.m, a.n, a.row_ptr, a.col_idx, np.abs(a.vals))


def test_cpp_host_layer(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/host_layer_driver is not built (run __graft_entry__.build())")
    a = gen.rmat(11, seed=21)
    b = gen.uniform(a.n, 900, 6, seed=22)
    c0 = gen.uniform(a.m, 900, 3, seed=23)
    for name, x in (("a", a), ("b", b), ("c0", c0)):
        write_binary(x, str(tmp_path / f"{name}.bin"))
    rng = np.random.default_rng(0)
    m, n, k = 130, 70, 90
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    C = rng.standard_normal((m, n)).astype(np.float32)
    with open(tmp_path / "dense.bin", "wb") as f:
        np.array([m, n, k], np.int64).tofile(f)
        A.tofile(f), B.tofile(f), C.tofile(f)
    sn, ld = 40, 48
    SB = rng.standard_normal((a.n, ld)).astype(np.float32)
    SC = rng.standard_normal((a.m, ld)).astype(np.float32)
    with open(tmp_path / "spmm_in.bin", "wb") as f:
        np.array([sn, ld], np.int64).tofile(f)
        SB.tofile(f), SC.tofile(f)
    print("[cpp] running", DRIVER, flush=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    def check_sparse(got, ref, ref64, absum, bound):
        np.testing.assert_array_equal(got.row_ptr, ref.row_ptr)
        np.testing.assert_array_equal(got.col_idx, ref.col_idx)
        assert np.all(np.abs(got.vals - ref64.vals) <= bound * absum.vals + 1e-30)

    u = 2.0 ** -24
    check_sparse(read_binary(str(tmp_path / "dot.bin")), O.spgemm(a, b), O.spgemm_f64(a, b),
                 O.spgemm_f64(_abs(a), _abs(b)), u)
    ref = O.spgemm(a, b, c0)
    ref64 = O.spgemm_f64(a, b)
    # c0 + a*b on the reference's structure, values within two roundings of the float64 sum
    d64 = ref64.to_dense() + c0.to_dense()
    absd = O.spgemm_f64(_abs(a), _abs(b)).to_dense() + np.abs(c0.to_dense())
    for fname in ("gemm.bin", "acc.bin"):
        g = read_binary(str(tmp_path / fname))
        np.testing.assert_array_equal(g.row_ptr, ref.row_ptr)
        np.testing.assert_array_equal(g.col_idx, ref.col_idx)
        rows = np.repeat(np.arange(g.m), g.row_nnz())
        assert np.all(np.abs(g.vals - d64[rows, g.col_idx]) <= 2 * u * absd[rows, g.col_idx] + 1e-30)
    dense64 = C.astype(np.float64) + O.gemm_f64(A, B)
    S = np.abs(C) + O.gemm_f64(np.abs(A), np.abs(B))
    for fname in ("gemm_out.bin", "gemm_cm.bin"):
        got = np.fromfile(tmp_path / fname, np.float32).reshape(m, n)
        assert np.all(np.abs(got - dense64) <= 2e-6 * S)
    got = np.fromfile(tmp_path / "spmm_out.bin", np.float32).reshape(a.m, ld)
    exp = SC[:, :sn].astype(np.float64) + O.spmm_f64(a, np.ascontiguousarray(SB[:, :sn]))
    mag = np.abs(SC[:, :sn]) + O.spmm_f64(_abs(a), np.ascontiguousarray(np.abs(SB[:, :sn])))
    assert np.all(np.abs(got[:, :sn] - exp) <= 1e-5 * mag + 1e-30)
    np.testing.assert_array_equal(got[:, sn:], SC[:, sn:])
