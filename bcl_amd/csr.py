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


def grid_tile_shape(m: int, n: int, grid_m: int, grid_n: int):
    """BlockCustom with Tile::div (bcl/containers/detail/Blocking.hpp:223-232): ceil(m/grid_m), ceil(n/grid_n)."""
    return (m + grid_m - 1) // grid_m, (n + grid_n - 1) // grid_n


def stage_pairs(grid: int, k: int):
    """The tile products of SUMMA stage k on a grid x grid blocking: (i, j) -> A_ik * B_kj.

    In BCL::gemm (bcl/containers/algorithms/spgemm.hpp:200-224) the owner of C(i, j) visits k in the staggered
    order (k_ + i + j) % K; gathering every owner's product for one k gives the grid x grid batch.
    """
    return [(i, j) for i in range(grid) for j in range(grid)]
