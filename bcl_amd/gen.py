"""Synthetic workloads for the local tile multiply (host C++ generators in csrc/gen.cpp, bound with ctypes).

Deterministic in (seed, sizes) and thread-count independent; see csrc/gen.cpp for the exact distributions.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .csr import CSR, grid_tile_shape

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libbclx_gen.so")
_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built: run __graft_entry__.build() (or python -m bcl_amd.build)")
        L = ctypes.CDLL(_LIB_PATH)
        L.bclx_gen_uniform.restype = ctypes.c_int
        L.bclx_gen_uniform.argtypes = [_I64, _I64, ctypes.c_int, ctypes.c_uint64, _P, _P, _P]
        L.bclx_gen_banded.restype = _I64
        L.bclx_gen_banded.argtypes = [_I64, _I64, _I64, ctypes.c_int, ctypes.c_uint64, _P, _P, _P]
        L.bclx_gen_rmat.restype = _I64
        L.bclx_gen_rmat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_uint64, _P, _P, _P]
        L.bclx_gen_laplacian2d.restype = _I64
        L.bclx_gen_laplacian2d.argtypes = [_I64, _P, _P, _P]
        L.bclx_csr_grid_counts.restype = None
        L.bclx_csr_grid_counts.argtypes = [_I64, _I64, _P, _P, _I64, _I64, _I64, _I64, _P]
        L.bclx_csr_tile_fp32.restype = ctypes.c_int
        L.bclx_csr_tile_fp32.argtypes = [_I64, _I64, _P, _P, _P, _I64, _I64, _I64, _I64, _P, _P, _P]
        _lib = L
    return _lib


def uniform(m: int, n: int, per_row: int = 8, seed: int = 0xB0C1) -> CSR:
    rp = np.zeros(m + 1, np.int64)
    col = np.zeros(m * per_row, np.int32)
    val = np.zeros(m * per_row, np.float32)
    if lib().bclx_gen_uniform(m, n, per_row, seed, _ptr(rp), _ptr(col), _ptr(val)) != 0:
        raise ValueError("bad uniform arguments")
    return CSR(m, n, rp, col, val)


def banded(m: int, n: int, bw: int = 64, max_per_row: int = 16, seed: int = 7) -> CSR:
    rp = np.zeros(m + 1, np.int64)
    col = np.zeros(m * max_per_row, np.int32)
    val = np.zeros(m * max_per_row, np.float32)
    nnz = lib().bclx_gen_banded(m, n, bw, max_per_row, seed, _ptr(rp), _ptr(col), _ptr(val))
    if nnz < 0:
        raise ValueError("bad banded arguments")
    return CSR(m, n, rp, col[:nnz].copy(), val[:nnz].copy())


RMAT_ABC = (0.57, 0.19, 0.19)


def rmat(scale: int, ef: int = 16, abc=RMAT_ABC, seed: int = 1234) -> CSR:
    n = 1 << scale
    rp = np.zeros(n + 1, np.int64)
    col = np.zeros(n * ef, np.int32)
    val = np.zeros(n * ef, np.float32)
    nnz = lib().bclx_gen_rmat(scale, ef, abc[0], abc[1], abc[2], seed, _ptr(rp), _ptr(col), _ptr(val))
    if nnz < 0:
        raise ValueError("bad R-MAT arguments")
    return CSR(n, n, rp, col[:nnz].copy(), val[:nnz].copy())


def laplacian2d(g: int) -> CSR:
    n = g * g
    nnz = 5 * n - 4 * g
    rp = np.zeros(n + 1, np.int64)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    got = lib().bclx_gen_laplacian2d(g, _ptr(rp), _ptr(col), _ptr(val))
    assert got == nnz
    return CSR(n, n, rp, col, val)


def split_grid(a: CSR, grid_m: int, grid_n: int):
    """Cut ``a`` into a grid of tiles (CSRMatrix::get_slice_impl_ semantics). Returns {(I, J): CSR}."""
    assert a.vals.dtype == np.float32
    tm, tn = grid_tile_shape(a.m, a.n, grid_m, grid_n)
    counts = np.zeros(grid_m * grid_n, np.int64)
    rp = np.ascontiguousarray(a.row_ptr, np.int64)
    lib().bclx_csr_grid_counts(a.m, a.n, _ptr(rp), _ptr(a.col_idx), tm, tn, grid_m, grid_n, _ptr(counts))
    tiles = {}
    for I in range(grid_m):
        for J in range(grid_n):
            rows = min(tm, a.m - I * tm)
            cols = min(tn, a.n - J * tn)
            nnz = int(counts[I * grid_n + J])
            trp = np.zeros(rows + 1, np.int32)
            tcol = np.zeros(max(nnz, 1), np.int32)
            tval = np.zeros(max(nnz, 1), np.float32)
            rc = lib().bclx_csr_tile_fp32(a.m, a.n, _ptr(rp), _ptr(a.col_idx), _ptr(a.vals), tm, tn, I, J,
                                          _ptr(trp), _ptr(tcol), _ptr(tval))
            assert rc == 0 and trp[-1] == nnz
            tiles[(I, J)] = CSR(rows, cols, trp.astype(np.int64), tcol[:nnz], tval[:nnz])
    return tiles
