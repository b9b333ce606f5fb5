"""The C++ host layer (include/bclx/bclx.hpp) driven like BCL code: CSRMatrix::dot, gemm (c += a * b),
SparseHashAccumulator, spmm_wrapper and cblas_gemm_wrapper_ (row- and column-major), through files in BCL's binary
CSR format; CSRMatrix(fname) of a MatrixMarket file and write_MatrixMarket, CSRMatrixMemoryMapped::get_slice_impl_,
split_grid and the DMatrix tile arithmetic. Results are checked against the oracle with the same bounds as the C-ABI
tests, and against the reference's own readers and slicers where oracle/_ref is built."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
from bcl_amd import gen
from bcl_amd.csr import CSR, get_slice, read_binary, read_matrix_market, write_binary, write_matrix_market

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
        write_binary(x, str(tmp_path / f"{name}.binary"))
    write_matrix_market(a, str(tmp_path / "a.mtx"))
    em, en, eld = 37, 50, 56
    X = np.random.default_rng(5).standard_normal((em, eld))
    Y = np.random.default_rng(6).standard_normal((em, eld)) + 3.0
    with open(tmp_path / "ew.bin", "wb") as f:
        np.array([em, en, eld], np.int64).tofile(f)
        X.tofile(f), Y.tofile(f)
    rng = np.random.default_rng(0)
    m, n, k = 130, 70, 90
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    C = rng.standard_normal((m, n)).astype(np.float32)
    with open(tmp_path / "dense.bin", "wb") as f:
        np.array([m, n, k], np.int64).tofile(f)
        A.tofile(f), B.tofile(f), C.tofile(f)
    m6, n6, k6 = 70, 90, 110
    A6, B6, C6 = (rng.standard_normal(sh) for sh in ((m6, k6), (k6, n6), (m6, n6)))
    with open(tmp_path / "dense64.bin", "wb") as f:
        np.array([m6, n6, k6], np.int64).tofile(f)
        A6.tofile(f), B6.tofile(f), C6.tofile(f)
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
    dot = read_binary(str(tmp_path / "dot.bin"))
    for other, same_as in (("dot_alloc.bin", "dot.bin"), ("dot_like.bin", "dot.bin"), ("gemm_like.bin", "gemm.bin"),
                           ("acc_like.bin", "acc.bin")):
        x, y = read_binary(str(tmp_path / other)), read_binary(str(tmp_path / same_as))
        np.testing.assert_array_equal(x.row_ptr, y.row_ptr)
        np.testing.assert_array_equal(x.col_idx, y.col_idx)
        np.testing.assert_array_equal(x.vals, y.vals)
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

    # fp64 overloads
    g6 = np.fromfile(tmp_path / "gemm64_out.bin", np.float64).reshape(2, m6, n6)
    S6 = np.abs(C6) + np.abs(A6) @ np.abs(B6)
    for q in range(2):
        assert np.all(np.abs(g6[q] - (C6 + A6 @ B6)) <= 2.0 * (k6 + 2) * 2.0 ** -53 * S6)
    s6 = np.fromfile(tmp_path / "spmm64_out.bin", np.float64).reshape(a.m, ld)
    import scipy.sparse
    a64 = scipy.sparse.csr_matrix((a.vals.astype(np.float64), a.col_idx, a.row_ptr), shape=(a.m, a.n))
    exp6 = SC[:, :sn].astype(np.float64) + a64 @ SB[:, :sn].astype(np.float64)
    mag6 = np.abs(SC[:, :sn]) + abs(a64) @ np.abs(SB[:, :sn]).astype(np.float64)
    L6 = np.diff(a.row_ptr)[:, None] + 2
    assert np.all(np.abs(s6[:, :sn] - exp6) <= 2 * L6 * 2.0 ** -53 * mag6 + 1e-300)
    np.testing.assert_array_equal(s6[:, sn:], SC[:, sn:].astype(np.float64))

    # files and tiles
    mtx = read_binary(str(tmp_path / "mtx.binary"))
    want = read_matrix_market(str(tmp_path / "a.mtx"))
    for x, y in ((mtx, want), (read_binary(str(tmp_path / "slice.binary")), get_slice(a, 100, 700, 50, 1500))):
        np.testing.assert_array_equal(x.row_ptr, y.row_ptr)
        np.testing.assert_array_equal(x.col_idx, y.col_idx)
        np.testing.assert_array_equal(x.vals, y.vals)
    assert open(tmp_path / "mtx_out.mtx", "rb").read() == open(tmp_path / "a.mtx", "rb").read()
    if O.have_ref():
        r = O.ref_mmap_slice(str(tmp_path / "a.binary"), 100, 700, 50, 1500)
        np.testing.assert_array_equal(read_binary(str(tmp_path / "slice.binary")).col_idx, r.col_idx)
    gm, gn = -(-a.m // 300), -(-a.n // 500)
    for t in range(gm * gn):
        i, j = divmod(t, gn)
        x, y = read_binary(str(tmp_path / f"split_{t}.binary")), get_slice(a, i * 300, (i + 1) * 300, j * 500,
                                                                           (j + 1) * 500)
        np.testing.assert_array_equal(x.row_ptr, y.row_ptr)
        np.testing.assert_array_equal(x.col_idx, y.col_idx)
        np.testing.assert_array_equal(x.vals, y.vals)
    ew = np.fromfile(tmp_path / "ew_out.bin", np.float64).reshape(5, em, eld)
    for q, want_ew in enumerate((X + Y, X - Y, X * Y, X / Y, X + Y)):
        np.testing.assert_array_equal(ew[q][:, :en], want_ew[:, :en])
    assert (ew[0][:, en:] == -1.0).all() and (ew[4][:, en:] == X[:, en:]).all()
