"""GPU cross-stage accumulation (bclx_accumulate_tiles) against the reference's SparseHashAccumulator.

Structure must equal the reference's get_matrix output (golden fixture) bit for bit. Values are summed in float64 and
rounded once: each must be within one fp32 rounding of the float64 sum, hence within 2 ulps of the reference's fp32
left-to-right sum over 4 tiles. The single-shot invariant: summing a C tile's partial products over all SUMMA stages
equals multiplying the whole block row by the whole block column.
"""
import os

import numpy as np
import pytest

import oracle_py as O
from bcl_amd import gen
from bcl_amd.csr import CSR
from bcl_amd.device import DeviceCSR, accumulate, spgemm

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_csr(d, prefix):
    m, n = (int(x) for x in d[f"{prefix}_shape"])
    return CSR(m, n, d[f"{prefix}_row_ptr"].astype(np.int64), d[f"{prefix}_col_idx"], d[f"{prefix}_vals"])


def test_golden_accumulate(handle):
    d = np.load(os.path.join(GOLDEN, "accumulate4.npz"), allow_pickle=False)
    tiles = [load_csr(d, f"t{t}") for t in range(4)]
    ref = load_csr(d, "c")
    c = accumulate(handle, [DeviceCSR.from_host(t) for t in tiles]).to_host()
    np.testing.assert_array_equal(c.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(c.col_idx, ref.col_idx)
    exact = np.zeros(c.nnz, np.float64)
    dense = sum(t.to_dense() for t in tiles)
    rows = np.repeat(np.arange(c.m), c.row_nnz())
    exact = dense[rows, c.col_idx]
    assert np.all(np.abs(c.vals - exact) <= 2.0 ** -24 * np.abs(exact) + 1e-30)
    ulp = np.abs(c.vals.view(np.int32).astype(np.int64) - ref.vals.view(np.int32).astype(np.int64))
    assert ulp.max() <= 2


def test_single_shot_invariant(handle):
    a = gen.rmat(12, seed=9)
    g = 4
    tiles = gen.split_grid(a, g, g)
    i, j = 1, 2
    parts = []
    for k in range(g):
        parts.append(spgemm(handle, DeviceCSR.from_host(tiles[(i, k)]), DeviceCSR.from_host(tiles[(k, j)])))
    c = accumulate(handle, parts).to_host()
    # whole block row i times whole block column j
    tm = (a.m + g - 1) // g
    rows = slice(i * tm, (i + 1) * tm)
    row_blk = CSR(tm, a.n, (a.row_ptr[rows.start:rows.stop + 1] - a.row_ptr[rows.start]).copy(),
                  a.col_idx[a.row_ptr[rows.start]:a.row_ptr[rows.stop]].copy(),
                  a.vals[a.row_ptr[rows.start]:a.row_ptr[rows.stop]].copy())
    col_blk = gen.split_grid(a, 1, g)[(0, j)]
    ref = O.spgemm(row_blk, col_blk)
    np.testing.assert_array_equal(c.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(c.col_idx, ref.col_idx)
    c64 = O.spgemm_f64(row_blk, col_blk)
    absum = O.spgemm_f64(CSR(row_blk.m, row_blk.n, row_blk.row_ptr, row_blk.col_idx, np.abs(row_blk.vals)),
                         CSR(col_blk.m, col_blk.n, col_blk.row_ptr, col_blk.col_idx, np.abs(col_blk.vals)))
    # two roundings: each partial product, then the sum
    assert np.all(np.abs(c.vals - c64.vals) <= 2.0 ** -23 * absum.vals + 1e-30)


def test_empty_and_fp64(handle):
    t = [CSR.empty(40, 30, np.float64) for _ in range(3)]
    c = accumulate(handle, [DeviceCSR.from_host(x) for x in t]).to_host()
    assert c.nnz == 0 and np.all(c.row_ptr == 0)
    lap = gen.laplacian2d(10)
    c = accumulate(handle, [DeviceCSR.from_host(lap), DeviceCSR.from_host(lap)]).to_host()
    np.testing.assert_array_equal(c.col_idx, lap.col_idx)
    np.testing.assert_array_equal(c.vals, 2 * lap.vals)
