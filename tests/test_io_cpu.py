"""CPU checks of the file formats either side of the tile multiply (no GPU).

- MatrixMarket reading and writing through libbclx (bclx_mm_read / bclx_mm_write / bclx_mm_info) against the
  reference's own CSRMatrix::read_MatrixMarket / write_MatrixMarket / matrix_io::matrix_market_info
  (bcl/containers/sequential/CSRMatrix.hpp:369-411, :532-548; matrix_io.hpp:68-98, :221-357), compiled from the
  reference sources into oracle/_ref: identical CSR (structure and value bits) on general, symmetric, pattern and
  integer files with comments and repeated entries, 0- and 1-based; the same errors on malformed files; byte-identical
  written files. The committed fixtures tests/golden/mm_*.npz hold the reference's results for when oracle/_ref is
  absent.
- A file too large for the reference's quadratic reader, read in linear time and checked against scipy.io.mmread
  (a file without repeated entries, where the two readers' semantics agree).
- The memory-mapped binary CSR loader and slicing against CSRMatrixMemoryMapped::get_slice_impl_ and
  CSRMatrix::get_slice_impl_ (CSRMatrixMemoryMapped.hpp:55-145, CSRMatrix.hpp:238-276).
"""
import os
import time

import numpy as np
import pytest

import oracle_py as O
from bcl_amd import gen
from bcl_amd._lib import BclxError
from bcl_amd.csr import (CSR, get_slice, matrix_market_info, read_binary_mmap, read_matrix_market, write_binary,
                         write_matrix_market)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
need_ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref (the reference build) is not present")


def mm_text(m, n, entries, field="real", symmetry="general", one_indexed=True, comments=2, fmt="{:.9g}"):
    """A MatrixMarket coordinate file: entries is a list of (i, j, v) with 0-based indices, written in order."""
    base = 1 if one_indexed else 0
    lines = [f"%%MatrixMarket matrix coordinate {field} {symmetry}"]
    lines += [f"% comment {c}" for c in range(comments)]
    lines.append(f"{m} {n} {len(entries)}")
    for i, j, v in entries:
        if field == "pattern":
            lines.append(f"{i + base} {j + base}")
        elif field == "integer":
            lines.append(f"{i + base} {j + base} {int(v)}")
        else:
            lines.append(f"{i + base} {j + base} " + fmt.format(v))
    return "\n".join(lines) + "\n"


def random_entries(rng, m, n, count, dup_frac=0.2, lower=False):
    ii = rng.integers(0, m, count)
    jj = rng.integers(0, n, count)
    if lower:  # symmetric files store one triangle
        ii, jj = np.maximum(ii, jj), np.minimum(ii, jj)
    vv = rng.standard_normal(count)
    ent = list(zip(ii.tolist(), jj.tolist(), vv.tolist()))
    ndup = int(dup_frac * count)
    for q in rng.integers(0, count, ndup):  # repeated (i, j) with other values, later in the file
        i, j, _ = ent[int(q)]
        ent.append((i, j, float(rng.standard_normal())))
    return ent


CASES = {
    "general_real": dict(m=300, n=200, field="real", symmetry="general", one=True, count=2500),
    "symmetric_real": dict(m=250, n=250, field="real", symmetry="symmetric", one=True, count=2000),
    "pattern_general": dict(m=180, n=400, field="pattern", symmetry="general", one=True, count=1500),
    "integer_symmetric": dict(m=128, n=128, field="integer", symmetry="symmetric", one=True, count=900),
    "zero_indexed": dict(m=64, n=90, field="real", symmetry="general", one=False, count=700),
    "empty_rows": dict(m=1000, n=50, field="real", symmetry="general", one=True, count=120),
}


def write_case(tmp_path, name, seed=0):
    c = CASES[name]
    rng = np.random.default_rng(seed + len(name))
    ent = random_entries(rng, c["m"], c["n"], c["count"], lower=c["symmetry"] == "symmetric")
    if c["field"] == "integer":
        ent = [(i, j, float(rng.integers(-50, 50))) for i, j, _ in ent]
    path = os.path.join(tmp_path, name + ".mtx")
    with open(path, "w") as f:
        f.write(mm_text(c["m"], c["n"], ent, c["field"], c["symmetry"], c["one"]))
    return path, c


def assert_same(a: CSR, b: CSR):
    assert (a.m, a.n, a.nnz) == (b.m, b.n, b.nnz)
    np.testing.assert_array_equal(a.row_ptr, b.row_ptr)
    np.testing.assert_array_equal(a.col_idx, b.col_idx)
    assert a.vals.dtype == b.vals.dtype
    np.testing.assert_array_equal(a.vals.view(np.uint8), b.vals.view(np.uint8))  # bit for bit


@need_ref
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mm_read_matches_reference(tmp_path, name, dtype):
    path, c = write_case(str(tmp_path), name)
    got = read_matrix_market(path, one_indexed=c["one"], dtype=dtype)
    got.check()
    ref = O.ref_mm_read(path, one_indexed=c["one"], dtype=dtype)
    assert_same(got, ref)
    assert matrix_market_info(path) == O.ref_mm_info(path)


def test_mm_read_golden(tmp_path):
    """The reference's results on the committed inputs (written by tests/golden/make_golden.py)."""
    d = np.load(os.path.join(GOLDEN, "mm_read.npz"))
    names = [k[:-len("_text")] for k in d.files if k.endswith("_text")]
    assert names
    for nm in names:
        path = os.path.join(str(tmp_path), nm + ".mtx")
        with open(path, "wb") as f:
            f.write(d[nm + "_text"].tobytes())
        got = read_matrix_market(path, one_indexed=bool(d[nm + "_one"][0]))
        want = CSR(int(d[nm + "_shape"][0]), int(d[nm + "_shape"][1]), d[nm + "_row_ptr"], d[nm + "_col_idx"],
                   d[nm + "_vals"])
        assert_same(got, want)


def test_mm_first_occurrence_wins_and_mirrors(tmp_path):
    """Closed-form case: repeated entries keep the first value; a symmetric off-diagonal entry stands for both
    triangles, and its mirror is inserted right after it (so a later explicit (j, i) is dropped)."""
    path = os.path.join(str(tmp_path), "s.mtx")
    with open(path, "w") as f:
        f.write(mm_text(3, 3, [(1, 0, 2.0), (2, 2, 5.0), (1, 0, 7.0), (0, 1, 9.0), (2, 1, -1.0)],
                        symmetry="symmetric"))
    a = read_matrix_market(path)
    np.testing.assert_array_equal(a.to_dense(), [[0, 2, 0], [2, 0, -1], [0, -1, 5]])


BAD = {
    "banner": ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "could not be parsed as a Matrix Market"),
    "type": ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1.0\n", "unsupported matrix type"),
    "bounds": ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n3 1 2.0\n", "out of bounds"),
    "count": ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.0\n2 2 2.0\n", "more nonzeros than reported"),
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_mm_read_errors(tmp_path, case):
    text, msg = BAD[case]
    path = os.path.join(str(tmp_path), case + ".mtx")
    with open(path, "w") as f:
        f.write(text)
    with pytest.raises(BclxError) as e:
        read_matrix_market(path)
    assert e.value.status == 1 and msg in str(e.value)
    if O.have_ref():
        with pytest.raises(O.RefError) as r:
            O.ref_mm_read(path)
        assert msg in str(r.value)


@need_ref
@pytest.mark.parametrize("name", ["general_real", "empty_rows"])
def test_mm_write_matches_reference_bytes(tmp_path, name):
    path, c = write_case(str(tmp_path), name)
    a = read_matrix_market(path, one_indexed=c["one"])
    ours, theirs = os.path.join(str(tmp_path), "ours.mtx"), os.path.join(str(tmp_path), "ref.mtx")
    write_matrix_market(a, ours)
    O.ref_mm_write(theirs, a)
    assert open(ours, "rb").read() == open(theirs, "rb").read()
    back = read_matrix_market(ours)  # "%f" keeps 6 decimals
    np.testing.assert_array_equal(back.col_idx, a.col_idx)
    np.testing.assert_allclose(back.vals, a.vals, atol=5e-7)


def test_mm_read_large_is_linear_time(tmp_path):
    """1.8 M entries (no repeats) in random order, read in well under a second, against scipy's reader."""
    import scipy.io
    import scipy.sparse
    a = gen.rmat(16, ef=32, seed=5)
    path = os.path.join(str(tmp_path), "big.mtx")
    rows = np.repeat(np.arange(a.m), a.row_nnz())
    perm = np.random.default_rng(1).permutation(a.nnz)  # file order unrelated to the CSR order
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real general\n{a.m} {a.n} {a.nnz}\n")
        body = np.stack([rows[perm] + 1, a.col_idx[perm].astype(np.int64) + 1], 1)
        vals = a.vals[perm]
        np.savetxt(f, np.column_stack([body, vals]), fmt=["%d", "%d", "%.9g"])
    t = time.perf_counter()
    got = read_matrix_market(path)
    dt = time.perf_counter() - t
    got.check()
    sp = scipy.sparse.csr_matrix(scipy.io.mmread(path))
    sp.sort_indices()
    np.testing.assert_array_equal(got.row_ptr, sp.indptr)
    np.testing.assert_array_equal(got.col_idx, sp.indices)
    np.testing.assert_array_equal(got.vals, sp.data.astype(np.float32))
    assert a.nnz > 1_500_000 and dt < 20.0, dt


@need_ref
def test_mmap_slices_match_reference(tmp_path):
    a = gen.rmat(12, seed=8)
    path = os.path.join(str(tmp_path), "a.binary")
    write_binary(a, path)
    mm = read_binary_mmap(path)
    assert isinstance(mm.col_idx, np.memmap)
    for (i0, i1, j0, j1) in [(0, a.m, 0, a.n), (100, 900, 1000, 3000), (4000, 5000, 0, 10), (7, 7, 0, a.n),
                             (a.m - 3, a.m + 50, a.n - 100, a.n + 100)]:
        got = get_slice(mm, i0, i1, j0, j1)
        assert_same(got, O.ref_mmap_slice(path, i0, i1, j0, j1))
        assert_same(get_slice(a, i0, i1, j0, j1), O.ref_slice(a, i0, i1, j0, j1))


def test_mmap_matches_read_binary(tmp_path):
    a = gen.rmat(11, seed=2)
    path = os.path.join(str(tmp_path), "a.binary")
    write_binary(a, path)
    mm = read_binary_mmap(path)
    assert (mm.m, mm.n, mm.nnz) == (a.m, a.n, a.nnz)
    np.testing.assert_array_equal(np.asarray(mm.row_ptr), a.row_ptr)
    np.testing.assert_array_equal(np.asarray(mm.col_idx), a.col_idx)
    np.testing.assert_array_equal(np.asarray(mm.vals), a.vals)
