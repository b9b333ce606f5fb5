"""CPU checks of the boundary and the host logic (no GPU, no compute calls into the HIP library).

- libbclx.so loads and exports every entry point include/bclx.h declares (and nothing named bclx_ is missing
  from the Python binding's signature table)
- the synthetic generators are deterministic and produce the documented shapes
- tile cutting follows CSRMatrix::get_slice_impl_ and BlockCustom tile shapes
"""
import ctypes
import os
import re

import numpy as np
import pytest

from bcl_amd import _lib, gen
from bcl_amd.csr import CSR, grid_tile_shape

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "bclx.h")).read()
    return sorted(set(re.findall(r"BCLX_API\s+[\w\s\*]+?\b(bclx_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = header_symbols()
    assert len(syms) >= 20
    assert "bclx_spgemm_symbolic" in syms and "bclx_spmm" in syms and "bclx_gemm" in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(header_symbols()) == bound


def test_version_and_error_string_without_gpu():
    L = _lib.lib()
    assert b"bclx" in L.bclx_version()
    assert isinstance(L.bclx_last_error(), bytes)


def test_null_handle_is_rejected_without_touching_a_gpu():
    L = _lib.lib()
    assert L.bclx_synchronize(None) == 1  # BCLX_ERR_INVALID
    assert b"null handle" in L.bclx_last_error()


def test_uniform_generator():
    a = gen.uniform(500, 300, 8, seed=3)
    a.check()
    assert np.all(a.row_nnz() == 8)
    b = gen.uniform(500, 300, 8, seed=3)
    np.testing.assert_array_equal(a.col_idx, b.col_idx)
    np.testing.assert_array_equal(a.vals, b.vals)
    assert a.vals.min() >= -1 and a.vals.max() < 1
    # 24-bit values: exact in fp32, multiples of 2^-23
    assert np.all((a.vals.astype(np.float64) * 2 ** 23) % 1 == 0)


def test_rmat_generator_dedup_and_skew():
    a = gen.rmat(12, seed=1)
    a.check()
    assert a.nnz <= 16 * 4096
    keys = np.repeat(np.arange(a.m), a.row_nnz()).astype(np.int64) * a.n + a.col_idx
    assert len(np.unique(keys)) == a.nnz
    # quadrant (0,0) holds about a = 0.57 of the draws at the top level
    top = (np.repeat(np.arange(a.m), a.row_nnz()) < a.m // 2) & (a.col_idx < a.n // 2)
    assert 0.5 < top.mean() < 0.7
    b = gen.rmat(12, seed=1)
    np.testing.assert_array_equal(a.col_idx, b.col_idx)


def test_laplacian_generator():
    g = 7
    a = gen.laplacian2d(g)
    a.check()
    assert a.nnz == 5 * g * g - 4 * g
    d = a.to_dense()
    np.testing.assert_array_equal(d, d.T)
    assert np.all(np.diag(d) == 4)


def test_banded_generator():
    a = gen.banded(400, 400, bw=64, max_per_row=16, seed=2)
    a.check()
    rows = np.repeat(np.arange(a.m), a.row_nnz())
    assert np.all(np.abs(a.col_idx - rows) <= 64)
    assert a.row_nnz().min() >= 1 and a.row_nnz().max() <= 16


@pytest.mark.parametrize("grid", [(3, 4), (8, 8)])
def test_split_grid_slices(grid):
    a = gen.rmat(10, seed=2)
    gm, gn = grid
    tm, tn = grid_tile_shape(a.m, a.n, gm, gn)
    assert (tm, tn) == ((a.m + gm - 1) // gm, (a.n + gn - 1) // gn)
    tiles = gen.split_grid(a, gm, gn)
    d = a.to_dense()
    for (I, J), t in tiles.items():
        t.check()
        np.testing.assert_array_equal(t.to_dense(), d[I * tm:(I + 1) * tm, J * tn:(J + 1) * tn])
    assert sum(t.nnz for t in tiles.values()) == a.nnz


def test_stage_pairs_cover_grid():
    from bcl_amd.csr import stage_pairs
    assert len(stage_pairs(8, 3)) == 64


def test_binary_format_matches_reference(tmp_path):
    """BCL's binary CSR file: written here, read and rewritten by the reference's CSRMatrix::read_Binary /
    write_Binary, read back here unchanged."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    from bcl_amd.csr import read_binary, write_binary
    a = gen.rmat(9, seed=4)
    p1, p2 = str(tmp_path / "a.bin"), str(tmp_path / "b.bin")
    write_binary(a, p1)
    b = read_binary(p1)
    np.testing.assert_array_equal(b.row_ptr, a.row_ptr)
    np.testing.assert_array_equal(b.col_idx, a.col_idx)
    np.testing.assert_array_equal(b.vals, a.vals)
    if not O.have_ref():
        pytest.skip("reference build not present")
    assert O.ref().ref_binary_roundtrip(p1.encode(), p2.encode()) == a.nnz
    assert open(p1, "rb").read() == open(p2, "rb").read()
