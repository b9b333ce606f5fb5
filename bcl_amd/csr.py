"""Host-side CSR tiles (numpy), the in-memory layout of BCL's CSRMatrix.

A tile is 0-based CSR: ``row_ptr`` has m+1 entries starting at 0, ``col_idx`` holds int32 columns sorted and
unique within each row, ``vals`` holds fp32 or fp64 values (bcl/containers/sequential/CSRMatrix.hpp:32-53).
The reference stores ``row_ptr`` in the index type ``I``; here the host keeps int64 offsets so products whose
nnz exceeds 2^31 - 1 stay representable, and narrows to int32 on request.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class CSR:
    m: int
    n: int
    row_ptr: np.ndarray  # int64 [m+1]
    col_idx: np.ndarray  # int32 [nnz]
    vals: np.ndarray  # float32/float64 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1]) if self.m > 0 else 0

    @property
    def shape(self):
        return (self.m, self.n)

    def row_nnz(self) -> np.ndarray:
        return np.diff(self.row_ptr)

    def astype(self, dtype) -> "CSR":
        return CSR(self.m, self.n, self.row_ptr, self.col_idx, self.vals.astype(dtype))

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.m, self.n), dtype=np.float64)
        rows = np.repeat(np.arange(self.m), self.row_nnz())
        np.add.at(d, (rows, self.col_idx.astype(np.int64)), self.vals.astype(np.float64))
        return d

    def check(self) -> None:
        """Structural validity: monotone row_ptr, in-range, strictly increasing columns per row."""
        rp = self.row_ptr
        assert rp.shape == (self.m + 1,) and rp[0] == 0
        assert np.all(np.diff(rp) >= 0)
        assert self.col_idx.shape == (self.nnz,) and self.vals.shape == (self.nnz,)
        if self.nnz:
            assert self.col_idx.min() >= 0 and self.col_idx.max() < self.n
            d = np.diff(self.col_idx.astype(np.int64))
            starts = np.zeros(self.nnz, dtype=bool)
            starts[rp[:-1][np.diff(rp) > 0]] = True
            assert np.all((d > 0) | starts[1:]), "columns must be strictly increasing within each row"

    @staticmethod
    def from_dense(d: np.ndarray, dtype=np.float32) -> "CSR":
        m, n = d.shape
        mask = d != 0
        rp = np.zeros(m + 1, np.int64)
        rp[1:] = np.cumsum(mask.sum(axis=1))
        rows, cols = np.nonzero(mask)
        return CSR(m, n, rp, cols.astype(np.int32), d[rows, cols].astype(dtype))

    @staticmethod
    def empty(m: int, n: int, dtype=np.float32) -> "CSR":
        return CSR(m, n, np.zeros(m + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, dtype))


def write_binary(a: CSR, path: str, index_dtype=np.int32) -> None:
    """BCL's binary CSR file (CSRMatrix::write_Binary, bcl/containers/sequential/CSRMatrix.hpp:513-530): m, n, nnz
    as size_t, then vals, col_ind, row_ptr (row_ptr and col_ind in the index type)."""
    with open(path, "wb") as f:
        np.array([a.m, a.n, a.nnz], np.uint64).tofile(f)
        np.ascontiguousarray(a.vals).tofile(f)
        np.ascontiguousarray(a.col_idx, index_dtype).tofile(f)
        np.ascontiguousarray(a.row_ptr, index_dtype).tofile(f)


def read_binary(path: str, val_dtype=np.float32, index_dtype=np.int32) -> CSR:
    """Read BCL's binary CSR file (CSRMatrix::read_Binary, CSRMatrix.hpp:488-511)."""
    with open(path, "rb") as f:
        m, n, nnz = (int(x) for x in np.fromfile(f, np.uint64, 3))
        vals = np.fromfile(f, val_dtype, nnz)
        col = np.fromfile(f, index_dtype, nnz).astype(np.int32)
        rp = np.fromfile(f, index_dtype, m + 1).astype(np.int64)
    return CSR(m, n, rp, col, vals)


def read_binary_mmap(path: str, val_dtype=np.float32, index_dtype=np.int32) -> CSR:
    """BCL's binary CSR file mapped into memory without reading it (CSRMatrixMemoryMapped,
    bcl/containers/sequential/CSRMatrixMemoryMapped.hpp:55-90): the arrays are read-only views of the file, paged
    in as they are used (a slice or a device upload touches only what it needs). Column indices and row offsets keep
    the file's index type."""
    head = np.memmap(path, np.uint64, mode="r", shape=(3,))
    m, n, nnz = (int(x) for x in head)
    del head
    vsz, isz = np.dtype(val_dtype).itemsize, np.dtype(index_dtype).itemsize
    vals = np.memmap(path, val_dtype, mode="r", offset=24, shape=(nnz,)) if nnz else np.zeros(0, val_dtype)
    col = (np.memmap(path, index_dtype, mode="r", offset=24 + nnz * vsz, shape=(nnz,)) if nnz
           else np.zeros(0, index_dtype))
    rp = np.memmap(path, index_dtype, mode="r", offset=24 + nnz * (vsz + isz), shape=(m + 1,))
    return CSR(m, n, rp, col, vals)


def get_slice(a: CSR, imin: int, imax: int, jmin: int, jmax: int) -> CSR:
    """The submatrix [imin, imax) x [jmin, jmax) with columns shifted by jmin (CSRMatrix::get_slice_impl_,
    bcl/containers/sequential/CSRMatrix.hpp:238-276; CSRMatrixMemoryMapped::get_slice_impl_,
    CSRMatrixMemoryMapped.hpp:92-145): bounds are clipped to the matrix, entries keep their order within a row.
    Vectorised over the rows of the slice, so a slice of a memory-mapped file reads only those rows."""
    imax, jmax = min(imax, a.m), min(jmax, a.n)
    imin, jmin = max(imin, 0), max(jmin, 0)
    assert imin <= imax and jmin <= jmax
    lo, hi = int(a.row_ptr[imin]), int(a.row_ptr[imax])
    col = np.asarray(a.col_idx[lo:hi]).astype(np.int64)
    keep = (col >= jmin) & (col < jmax)
    rows = np.repeat(np.arange(imax - imin), np.diff(np.asarray(a.row_ptr[imin:imax + 1]).astype(np.int64)))
    rp = np.zeros(imax - imin + 1, np.int64)
    np.cumsum(np.bincount(rows[keep], minlength=imax - imin), out=rp[1:])
    return CSR(imax - imin, jmax - jmin, rp, (col[keep] - jmin).astype(np.int32), np.asarray(a.vals[lo:hi])[keep])


def matrix_market_info(path: str):
    """(m, n, stored entries) of a MatrixMarket file's size line (matrix_io::matrix_market_info,
    bcl/containers/sequential/matrix_io.hpp:68-98)."""
    import ctypes
    from ._lib import call
    v = [ctypes.c_int64() for _ in range(3)]
    call("bclx_mm_info", path.encode(), 1, *(ctypes.byref(x) for x in v))
    return tuple(int(x.value) for x in v)


def read_matrix_market(path: str, one_indexed: bool = True, dtype=np.float32) -> CSR:
    """A MatrixMarket coordinate file as CSR, sorted by (i, j), the first occurrence of a repeated entry kept and
    symmetric files mirrored (CSRMatrix::read_MatrixMarket over matrix_io::mmread,
    bcl/containers/sequential/CSRMatrix.hpp:369-411, matrix_io.hpp:221-357), read by libbclx in linear time.
    Raises BclxError with the reference's messages on malformed files."""
    import ctypes
    from ._lib import BCLX_F32, BCLX_F64, bclx_csr, call, lib
    f64 = np.dtype(dtype) == np.float64
    out = bclx_csr()
    call("bclx_mm_read", path.encode(), int(one_indexed), BCLX_F64 if f64 else BCLX_F32, 8, ctypes.byref(out))
    try:
        m, nnz = out.m, out.nnz
        rp = np.ctypeslib.as_array(ctypes.cast(out.row_ptr, ctypes.POINTER(ctypes.c_int64)), shape=(m + 1,)).copy()
        if nnz:
            col = np.ctypeslib.as_array(ctypes.cast(out.col_idx, ctypes.POINTER(ctypes.c_int32)), shape=(nnz,)).copy()
            ct = ctypes.c_double if f64 else ctypes.c_float
            val = np.ctypeslib.as_array(ctypes.cast(out.vals, ctypes.POINTER(ct)), shape=(nnz,)).copy()
        else:
            col, val = np.zeros(0, np.int32), np.zeros(0, np.float64 if f64 else np.float32)
    finally:
        for ptr in (out.row_ptr, out.col_idx, out.vals):
            lib().bclx_host_free(ptr)
    return CSR(out.m, out.n, rp, col, val)


def write_matrix_market(a: CSR, path: str) -> None:
    """CSRMatrix::write_MatrixMarket's format (bcl/containers/sequential/CSRMatrix.hpp:532-548), written by
    libbclx."""
    import ctypes
    from ._lib import BCLX_F32, BCLX_F64, bclx_csr, call
    rp = np.ascontiguousarray(a.row_ptr, np.int64)
    col = np.ascontiguousarray(a.col_idx, np.int32)
    f64 = a.vals.dtype == np.float64
    val = np.ascontiguousarray(a.vals, np.float64 if f64 else np.float32)
    d = bclx_csr(a.m, a.n, a.nnz, 8, BCLX_F64 if f64 else BCLX_F32, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    call("bclx_mm_write", path.encode(), ctypes.byref(d))


def grid_tile_shape(m: int, n: int, grid_m: int, grid_n: int):
    """BlockCustom with Tile::div (bcl/containers/detail/Blocking.hpp:223-232): ceil(m/grid_m), ceil(n/grid_n)."""
    return (m + grid_m - 1) // grid_m, (n + grid_n - 1) // grid_n


def stage_pairs(grid: int, k: int):
    """The tile products of SUMMA stage k on a grid x grid blocking: (i, j) -> A_ik * B_kj.

    In BCL::gemm (bcl/containers/algorithms/spgemm.hpp:200-224) the owner of C(i, j) visits k in the staggered
    order (k_ + i + j) % K; gathering every owner's product for one k gives the grid x grid batch.
    """
    return [(i, j) for i in range(grid) for j in range(grid)]
