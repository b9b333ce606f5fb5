"""Device-resident tiles and the Python face of the local tile multiply.

Device memory comes from torch (HBM tensors on cuda:N); every computation goes through libbclx.so's C ABI on
the handle's stream, which is torch's current stream when the handle is created. Host tiles are
bcl_amd.csr.CSR (the layout of BCL's CSRMatrix).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from . import _lib
from ._lib import BCLX_F32, BCLX_F64, bclx_csr, call, lib
from .csr import CSR

_DT = {torch.float32: BCLX_F32, torch.float64: BCLX_F64}


def _ptr(t: torch.Tensor | None):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


class Handle:
    """A bclx handle on one device, ordered on a HIP stream (torch's current stream by default)."""

    def __init__(self, device: int = 0, stream: torch.cuda.Stream | None = None):
        if not torch.cuda.is_available():
            raise RuntimeError("bclx needs a GPU (torch.cuda.is_available() is False)")
        self.device = device
        self.stream = stream or torch.cuda.current_stream(device)
        h = ctypes.c_void_p()
        call("bclx_create", device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h))
        self.h = h

    def close(self):
        if self.h:
            call("bclx_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        call("bclx_synchronize", self.h)


@dataclasses.dataclass
class DeviceCSR:
    m: int
    n: int
    row_ptr: torch.Tensor  # int32 or int64 [m+1]
    col_idx: torch.Tensor  # int32 [nnz]
    vals: torch.Tensor  # float32 / float64 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.col_idx.numel())

    def desc(self) -> bclx_csr:
        d = bclx_csr()
        d.m, d.n, d.nnz = self.m, self.n, self.nnz
        d.ptr_width = 8 if self.row_ptr.dtype == torch.int64 else 4
        d.dtype = _DT[self.vals.dtype]
        d.row_ptr = _ptr(self.row_ptr)
        d.col_idx = _ptr(self.col_idx)
        d.vals = _ptr(self.vals)
        return d

    @staticmethod
    def from_host(a: CSR, device="cuda", ptr_dtype=torch.int32) -> "DeviceCSR":
        rp = torch.from_numpy(np.ascontiguousarray(a.row_ptr)).to(device=device, dtype=ptr_dtype)
        col = torch.from_numpy(np.ascontiguousarray(a.col_idx, np.int32)).to(device)
        val = torch.from_numpy(np.ascontiguousarray(a.vals)).to(device)
        return DeviceCSR(a.m, a.n, rp, col, val)

    def to_host(self) -> CSR:
        return CSR(self.m, self.n, self.row_ptr.cpu().numpy().astype(np.int64), self.col_idx.cpu().numpy(),
                   self.vals.cpu().numpy())


def spgemm(h: Handle, A: DeviceCSR, B: DeviceCSR, ptr_dtype=torch.int64, return_rows: bool = False):
    """C = A * B through the two-phase C ABI (bclx_spgemm_symbolic, then bclx_spgemm_numeric)."""
    da, db = A.desc(), B.desc()
    plan = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    call("bclx_spgemm_symbolic", h.h, ctypes.byref(da), ctypes.byref(db), ctypes.byref(plan), ctypes.byref(nnz))
    try:
        dev = A.vals.device
        C = DeviceCSR(A.m, B.n, torch.empty(A.m + 1, dtype=ptr_dtype, device=dev),
                      torch.empty(nnz.value, dtype=torch.int32, device=dev),
                      torch.empty(nnz.value, dtype=A.vals.dtype, device=dev))
        dc = C.desc()
        dc.nnz = nnz.value
        call("bclx_spgemm_numeric", plan, ctypes.byref(dc))
        rows = None
        if return_rows:
            prod = np.zeros(A.m, np.int64)
            rn = np.zeros(A.m, np.int32)
            call("bclx_spgemm_plan_rows", plan, prod.ctypes.data_as(ctypes.c_void_p),
                 rn.ctypes.data_as(ctypes.c_void_p))
            rows = (prod, rn)
    finally:
        call("bclx_spgemm_plan_destroy", plan)
    return (C, rows) if return_rows else C


def spmm(h: Handle, A: DeviceCSR, B: torch.Tensor, C: torch.Tensor, n: int | None = None, alpha: float = 1.0,
         beta: float = 1.0) -> torch.Tensor:
    """C = beta * C + alpha * A @ B[:, :n] in place (BCL's spmm_wrapper is alpha = beta = 1). B row-major fp32 or
    bf16 (k x ldb), C row-major fp32 (m x ldc); or A, B and C all fp64 (bclx_spmm_f64)."""
    assert B.dim() == 2 and C.dim() == 2 and B.stride(1) == 1 and C.stride(1) == 1
    n = B.shape[1] if n is None else n
    d = A.desc()
    if C.dtype == torch.float64:  # the fp64 product: A, B and C all fp64
        assert B.dtype == torch.float64 and A.vals.dtype == torch.float64
        call("bclx_spmm_f64", h.h, ctypes.byref(d), _ptr(B), B.stride(0), _ptr(C), C.stride(0), n, alpha, beta)
        return C
    assert C.dtype == torch.float32 and B.dtype in (torch.float32, torch.bfloat16)
    bt = _lib.BCLX_F32 if B.dtype == torch.float32 else _lib.BCLX_BF16
    call("bclx_spmm", h.h, ctypes.byref(d), _ptr(B), bt, B.stride(0), _ptr(C), C.stride(0), n, alpha, beta)
    return C


def accumulate(h: Handle, tiles: list, ptr_dtype=torch.int64) -> DeviceCSR:
    """C = sum of partial tiles (SparseHashAccumulator::get_matrix on the GPU), two calls through the C ABI."""
    arr = (bclx_csr * len(tiles))(*[t.desc() for t in tiles])
    dev = tiles[0].vals.device
    C = DeviceCSR(tiles[0].m, tiles[0].n, torch.empty(tiles[0].m + 1, dtype=ptr_dtype, device=dev),
                  torch.empty(0, dtype=torch.int32, device=dev), torch.empty(0, dtype=tiles[0].vals.dtype, device=dev))
    d = C.desc()
    d.col_idx = None
    d.vals = None
    call("bclx_accumulate_tiles", h.h, len(tiles), arr, ctypes.byref(d))
    nnz = d.nnz
    if nnz == 0:
        return C  # the first call already wrote row_ptr
    C.col_idx = torch.empty(nnz, dtype=torch.int32, device=dev)
    C.vals = torch.empty(nnz, dtype=tiles[0].vals.dtype, device=dev)
    d = C.desc()
    d.nnz = nnz
    call("bclx_accumulate_tiles", h.h, len(tiles), arr, ctypes.byref(d))
    return C


def split_grid(h: Handle, A: DeviceCSR, tile_m: int, tile_n: int, ptr_dtype=torch.int32) -> dict:
    """The tiles of A on a grid of tile_m x tile_n tiles, cut on the GPU (SPMatrix's constructor, through
    CSRMatrix::get_slice_impl_): {(ti, tj): DeviceCSR}. Two calls through the C ABI: sizes, then the copy."""
    gm = max(1, -(-A.m // tile_m))
    gn = max(1, -(-A.n // tile_n))
    arr = (bclx_csr * (gm * gn))()
    da = A.desc()
    call("bclx_csr_split_grid", h.h, ctypes.byref(da), tile_m, tile_n, gm * gn, arr)
    dev = A.vals.device
    out = {}
    for t in range(gm * gn):
        d = arr[t]
        T = DeviceCSR(d.m, d.n, torch.empty(d.m + 1, dtype=ptr_dtype, device=dev),
                      torch.empty(d.nnz, dtype=torch.int32, device=dev), torch.empty(d.nnz, dtype=A.vals.dtype,
                                                                                      device=dev))
        arr[t] = T.desc()
        out[(t // gn, t % gn)] = T
    call("bclx_csr_split_grid", h.h, ctypes.byref(da), tile_m, tile_n, gm * gn, arr)
    return out


OP_ADD, OP_SUB, OP_MUL, OP_DIV = 0, 1, 2, 3


def dense_binary_op(h: Handle, op: int, A: torch.Tensor, B: torch.Tensor, C: torch.Tensor | None = None) -> torch.Tensor:
    """C = A op B elementwise on row-major tiles (DMatrix's +, -, *, /); C = A for the in-place form (+=)."""
    C = A if C is None else C
    assert A.dim() == B.dim() == C.dim() == 2 and A.shape == B.shape == C.shape and A.dtype == B.dtype == C.dtype
    assert A.dtype in _DT and all(t.stride(1) == 1 or t.shape[1] <= 1 for t in (A, B, C))
    m, n = A.shape
    ld = lambda t: max(t.stride(0), n)
    call("bclx_dense_binary_op", h.h, op, _DT[A.dtype], m, n, _ptr(A), ld(A), _ptr(B), ld(B), _ptr(C), ld(C))
    return C


GEMM_FP32, GEMM_BF16, GEMM_SPLIT_BF16 = 0, 1, 2


def gemm(h: Handle, A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, mode: int = GEMM_FP32, trans_a: bool = False,
         trans_b: bool = False, alpha: float = 1.0, beta: float = 1.0) -> torch.Tensor:
    """C = beta * C + alpha * op(A) @ op(B) in place, row-major fp32 (cblas_gemm_wrapper_ with RowMajor).
    mode: GEMM_FP32 (FP32 MFMA), GEMM_BF16 (bf16 inputs, fp32 accumulation), GEMM_SPLIT_BF16 (3-term split)."""
    for t in (A, B, C):
        assert t.dim() == 2 and t.dtype == torch.float32
        assert t.numel() == 0 or t.stride(1) == 1 or t.shape[1] <= 1, "rows must be contiguous"
    m, k = (A.shape[1], A.shape[0]) if trans_a else (A.shape[0], A.shape[1])
    kb, n = (B.shape[1], B.shape[0]) if trans_b else (B.shape[0], B.shape[1])
    assert k == kb and C.shape[0] >= m and C.shape[1] >= n
    lda = max(A.stride(0), A.shape[1])
    ldb = max(B.stride(0), B.shape[1])
    call("bclx_gemm", h.h, int(trans_a), int(trans_b), m, n, k, alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(C),
         C.stride(0), mode)
    return C


def gemm_f64(h: Handle, A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, trans_a: bool = False,
             trans_b: bool = False, alpha: float = 1.0, beta: float = 1.0) -> torch.Tensor:
    """C = beta * C + alpha * op(A) @ op(B) in place, row-major fp64 on the fp64 matrix cores (cblas_dgemm's place)."""
    for t in (A, B, C):
        assert t.dim() == 2 and t.dtype == torch.float64
        assert t.numel() == 0 or t.stride(1) == 1 or t.shape[1] <= 1, "rows must be contiguous"
    m, k = (A.shape[1], A.shape[0]) if trans_a else (A.shape[0], A.shape[1])
    kb, n = (B.shape[1], B.shape[0]) if trans_b else (B.shape[0], B.shape[1])
    assert k == kb and C.shape[0] >= m and C.shape[1] >= n
    call("bclx_gemm_f64", h.h, int(trans_a), int(trans_b), m, n, k, alpha, _ptr(A), max(A.stride(0), A.shape[1]),
         _ptr(B), max(B.stride(0), B.shape[1]), beta, _ptr(C), C.stride(0))
    return C


def gemm_f64_ref(h: Handle, A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """float64 A @ B on the GPU (fp32 inputs widened exactly): the error oracle for large dense products."""
    m, k = A.shape
    n = B.shape[1]
    C = torch.zeros((m, n), dtype=torch.float64, device=A.device)
    call("bclx_gemm_f64_ref", h.h, m, n, k, _ptr(A), A.stride(0), _ptr(B), B.stride(0), _ptr(C), C.stride(0))
    return C


def version() -> str:
    return lib().bclx_version().decode()


__all__ = ["Handle", "DeviceCSR", "spgemm", "spmm", "accumulate", "split_grid", "dense_binary_op", "gemm",
           "gemm_f64", "gemm_f64_ref", "version", "_lib"]
