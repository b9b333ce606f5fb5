"""CPU checks that pin the oracle (no GPU).

The oracle (oracle/oracle.c) restates the reference's CPU tile multiply. It is pinned by
  - the golden fixtures in tests/golden (outputs of the reference's own BCL::gemm / spmm_wrapper /
    SparseHashAccumulator compiled from its sources), bit for bit including fp32 value bits;
  - the reference build itself (oracle/_ref) on further seeded inputs, when present;
  - mathematics: brute-force dense products, scipy's boolean product on positive values, A * I = A, permutations,
    the closed form of the 2D Laplacian squared, B = ones for SpMM, fill_range inputs for GEMM, and finite
    differences for the MLP gradient.
Each check is chosen so that a plausible mistake (dropped term, transposed operand, wrong index, lost zero entry)
fails it.
"""
import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_py as O
from bcl_amd import gen
from bcl_amd.csr import CSR

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_csr(d, prefix):
    m, n = (int(x) for x in d[f"{prefix}_shape"])
    return CSR(m, n, d[f"{prefix}_row_ptr"].astype(np.int64), d[f"{prefix}_col_idx"], d[f"{prefix}_vals"])


def assert_csr_equal(c: CSR, ref: CSR, bits=True):
    assert (c.m, c.n, c.nnz) == (ref.m, ref.n, ref.nnz)
    np.testing.assert_array_equal(c.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(c.col_idx, ref.col_idx)
    if bits:
        u = np.uint32 if c.vals.dtype == np.float32 else np.uint64
        np.testing.assert_array_equal(c.vals.view(u), ref.vals.astype(c.vals.dtype).view(u))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "spgemm_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_spgemm_matches_golden(path):
    d = np.load(path, allow_pickle=False)
    a, b, ref = load_csr(d, "a"), load_csr(d, "b"), load_csr(d, "c")
    c0 = load_csr(d, "c0") if "c0_shape" in d else None
    assert_csr_equal(O.spgemm(a, b, c0), ref)


def test_spmm_matches_golden():
    d = np.load(os.path.join(GOLDEN, "spmm_rmat9.npz"), allow_pickle=False)
    a = load_csr(d, "a")
    n = int(d["n"][0])
    c = d["c0"].copy()
    b = d["b"]
    import ctypes
    arp, acol = np.ascontiguousarray(a.row_ptr, np.int64), np.ascontiguousarray(a.col_idx, np.int32)
    P = ctypes.c_void_p
    O.lib().oracle_spmm_f32(a.m, n, a.n, arp.ctypes.data_as(P), acol.ctypes.data_as(P),
                            a.vals.ctypes.data_as(P), b.ctypes.data_as(P), b.shape[1], c.ctypes.data_as(P),
                            c.shape[1])
    np.testing.assert_array_equal(c.view(np.uint32), d["c"].view(np.uint32))
    # columns past n are untouched (ld > n)
    np.testing.assert_array_equal(c[:, n:], d["c0"][:, n:])


def test_accumulate_matches_golden():
    d = np.load(os.path.join(GOLDEN, "accumulate4.npz"), allow_pickle=False)
    tiles = [load_csr(d, f"t{t}") for t in range(4)]
    ref = load_csr(d, "c")
    assert_csr_equal(O.accumulate(tiles, ref.m, ref.n, nthreads=2), ref)


@pytest.mark.skipif(not O.have_ref(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_spgemm_matches_reference_build(seed):
    a = gen.rmat(10, seed=seed)
    b = gen.uniform(a.n, 700, 5, seed=seed + 100)
    assert_csr_equal(O.spgemm(a, b), O.ref_spgemm(a, b))
    assert_csr_equal(O.spgemm(a, a), O.ref_spgemm(a, a))


def test_spgemm_brute_force_dense():
    rng = np.random.default_rng(0)
    for trial in range(20):
        m, k, n = rng.integers(1, 40, 3)
        da = np.where(rng.random((m, k)) < 0.2, rng.integers(-3, 4, (m, k)), 0).astype(np.float32)
        db = np.where(rng.random((k, n)) < 0.2, rng.integers(-3, 4, (k, n)), 0).astype(np.float32)
        a, b = CSR.from_dense(da), CSR.from_dense(db)
        c = O.spgemm(a, b)
        # pattern = boolean product of the patterns (zeros from cancellation stay)
        pat = ((da != 0).astype(np.int64) @ (db != 0).astype(np.int64)) > 0
        got = np.zeros((m, n), bool)
        rows = np.repeat(np.arange(m), c.row_nnz())
        got[rows, c.col_idx] = True
        np.testing.assert_array_equal(got, pat)
        np.testing.assert_array_equal(c.to_dense(), da.astype(np.float64) @ db.astype(np.float64))


def test_spgemm_structure_vs_scipy_positive_values():
    a = gen.rmat(11, seed=4)
    a.vals = np.abs(a.vals) + np.float32(0.5)  # no cancellation: scipy keeps every entry too
    c = O.spgemm(a, a)
    s = sp.csr_matrix((a.vals.astype(np.float64), a.col_idx, a.row_ptr), shape=a.shape)
    ss = (s @ s).tocsr()
    ss.sort_indices()
    np.testing.assert_array_equal(c.row_ptr, ss.indptr)
    np.testing.assert_array_equal(c.col_idx, ss.indices)
    np.testing.assert_allclose(c.vals, ss.data, rtol=1e-5)


def test_identity_and_permutation():
    a = gen.rmat(9, seed=5)
    eye = CSR(a.n, a.n, np.arange(a.n + 1, dtype=np.int64), np.arange(a.n, dtype=np.int32),
              np.ones(a.n, np.float32))
    assert_csr_equal(O.spgemm(a, eye), a)
    perm = np.random.default_rng(1).permutation(a.n).astype(np.int32)
    p = CSR(a.n, a.n, np.arange(a.n + 1, dtype=np.int64), perm, np.ones(a.n, np.float32))
    ap = O.spgemm(a, p)  # column j of A moves to perm[j]
    dense = a.to_dense()
    expect = np.zeros_like(dense)
    expect[:, perm] = dense
    np.testing.assert_array_equal(ap.to_dense(), expect)


def test_laplacian_squared_closed_form():
    """L = 4I - adjacency of the g x g grid. (L^2)_vv = 16 + deg(v); (L^2)_uv = -8 for grid neighbours,
    2 for diagonal neighbours (two 2-paths), 1 at grid distance 2 in a straight line (one 2-path)."""
    g = 9
    lap = gen.laplacian2d(g)
    c = O.spgemm(lap, lap)
    d = c.to_dense()
    for v in range(g * g):
        x, y = divmod(v, g)
        for u in range(g * g):
            ux, uy = divmod(u, g)
            dx, dy = abs(ux - x), abs(uy - y)
            deg = (x > 0) + (x < g - 1) + (y > 0) + (y < g - 1)
            if u == v:
                e = 16 + deg
            elif dx + dy == 1:
                e = -8
            elif dx == 1 and dy == 1:
                e = 2
            elif (dx == 2 and dy == 0) or (dx == 0 and dy == 2):
                e = 1
            else:
                e = 0
            assert d[v, u] == e, (v, u, d[v, u], e)
    # nnz: 13-point stencil clipped to the grid, no cancellation
    assert c.nnz == int((d != 0).sum())


def test_f64_oracle_vs_numpy():
    a = gen.uniform(200, 150, 7, seed=6)
    b = gen.banded(150, 180, bw=40, max_per_row=9, seed=7)
    c = O.spgemm_f64(a, b)
    np.testing.assert_allclose(c.to_dense(), a.to_dense() @ b.to_dense(), rtol=1e-14, atol=1e-14)


def test_rowhash_and_row_nnz_agree():
    a = gen.rmat(11, seed=8)
    ref = O.spgemm(a, a)
    rh = O.spgemm_rowhash(a, a, nthreads=4)
    np.testing.assert_array_equal(rh.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(rh.col_idx, ref.col_idx)
    np.testing.assert_allclose(rh.vals, ref.vals, rtol=1e-4, atol=1e-5)
    nnz, prod = O.row_nnz(a, a, 4)
    np.testing.assert_array_equal(nnz, np.diff(ref.row_ptr))
    blen = np.diff(a.row_ptr)
    expect_prod = np.add.reduceat(np.concatenate([blen[a.col_idx], [0]]), a.row_ptr[:-1]) * (np.diff(a.row_ptr) > 0)
    np.testing.assert_array_equal(prod, expect_prod)


def test_spmm_ones_closed_form_and_f64():
    a = gen.rmat(9, seed=9)
    b = np.ones((a.n, 16), np.float32)
    c = np.zeros((a.m, 16), np.float32)
    O.spmm(a, b, c)
    rowsum = np.add.reduceat(np.concatenate([a.vals.astype(np.float64), [0]]), a.row_ptr[:-1])
    rowsum[np.diff(a.row_ptr) == 0] = 0
    np.testing.assert_allclose(c, np.repeat(rowsum[:, None], 16, 1), rtol=1e-5, atol=1e-5)
    rng = np.random.default_rng(0)
    b = rng.standard_normal((a.n, 16)).astype(np.float32)
    c0 = rng.standard_normal((a.m, 16)).astype(np.float32)
    np.testing.assert_allclose(O.spmm_f64(a, b, c0), c0 + a.to_dense() @ b.astype(np.float64), rtol=1e-12,
                               atol=1e-12)
    bh = (b.view(np.uint32) >> 16).astype(np.uint16)  # truncated bf16 bits
    bb = (bh.astype(np.uint32) << 16).view(np.float32)
    np.testing.assert_allclose(O.spmm_f64(a, bh), a.to_dense() @ bb.astype(np.float64), rtol=1e-12, atol=1e-12)


def test_gemm_f64_oracle():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((70, 50)).astype(np.float32)
    b = rng.standard_normal((50, 90)).astype(np.float32)
    c0 = rng.standard_normal((70, 90))
    got = O.gemm_f64(a, b, c0)
    np.testing.assert_allclose(got, c0 + a.astype(np.float64) @ b.astype(np.float64), rtol=1e-13, atol=1e-12)
    part = O.gemm_f64(a, b, None, rows=(13, 29))
    np.testing.assert_allclose(part, (a.astype(np.float64) @ b.astype(np.float64))[13:29], rtol=1e-13, atol=1e-12)
    # fill_range closed form (bcl/containers/DMatrix.hpp:589-604): a_ij = (i*K + j) % 71, b_ij = (i*N + j) % 121
    M, N, K = 33, 27, 41
    fa = (np.arange(M * K) % 71).reshape(M, K).astype(np.float32)
    fb = (np.arange(K * N) % 121).reshape(K, N).astype(np.float32)
    exact = fa.astype(np.int64) @ fb.astype(np.int64)
    np.testing.assert_array_equal(O.gemm_f64(fa, fb), exact.astype(np.float64))


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1.5, size=(50, 16)).astype(np.float32)
    t = rng.uniform(0, 12, size=50).astype(np.float32)
    from bcl_amd import predictor
    params = predictor.init_params(seed=7)
    _, g64, _ = O.mlp(x, params, t)

    def fd(idx, h):
        pp = params.copy()
        pp[idx] += h
        pm = params.copy()
        pm[idx] -= h
        return (O.mlp(x, pp, t)[2] - O.mlp(x, pm, t)[2]) / (2 * h)

    kinks = 0
    for idx in rng.choice(params.size, 60, replace=False):
        f1, f2 = fd(idx, 1e-3), fd(idx, 2.5e-4)
        if abs(f1 - f2) > 1e-3 * max(1.0, abs(f1)):
            kinks += 1  # a ReLU switches inside the interval: the loss is not smooth there
            continue
        assert abs(f2 - g64[idx]) <= 1e-3 * max(1.0, abs(g64[idx])), (idx, f2, g64[idx])
    assert kinks <= 6


def test_planted_mistake_is_caught():
    """A plausible oracle bug (dropping exact-zero sums, as scipy does) must fail the golden check."""
    d = np.load(os.path.join(GOLDEN, "spgemm_int_cancel.npz"), allow_pickle=False)
    a, b, ref = load_csr(d, "a"), load_csr(d, "b"), load_csr(d, "c")
    assert (ref.vals == 0).any(), "fixture must contain explicit zeros"
    s = sp.csr_matrix((a.vals, a.col_idx, a.row_ptr), shape=a.shape)
    t = sp.csr_matrix((b.vals, b.col_idx, b.row_ptr), shape=b.shape)
    st = (s @ t).tocsr()
    st.eliminate_zeros()
    assert st.nnz != ref.nnz
