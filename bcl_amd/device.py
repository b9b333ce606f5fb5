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


def version() -> str:
    return lib().bclx_version().decode()


__all__ = ["Handle", "DeviceCSR", "spgemm", "version", "_lib"]
