"""ctypes binding of libbclx.so (the C ABI in include/bclx.h).

The library is required: there is no fallback. Loading fails loudly when it is not built, and every call that
returns a non-zero status raises BclxError with the library's message.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbclx.so")

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_F32 = ctypes.c_float

BCLX_F32, BCLX_F64, BCLX_BF16 = 0, 1, 2
MLP_NPARAM = 5313
MLP_NFEAT = 16


class BclxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"bclx status {status}: {msg}")
        self.status = status


class bclx_csr(ctypes.Structure):
    _fields_ = [("m", _I64), ("n", _I64), ("nnz", _I64), ("ptr_width", _I32), ("dtype", _I32),
                ("row_ptr", _P), ("col_idx", _P), ("vals", _P)]


class bclx_stage_opts(ctypes.Structure):
    _fields_ = [("mode", _I32), ("use_graph", _I32), ("margin", _F32), ("slack", _I32),
                ("h_mlp_params", _P), ("persistent_blocks", _I32), ("h_tuning", _P)]


# (name, restype, argtypes) for every entry point of include/bclx.h
SIGNATURES = [
    ("bclx_last_error", ctypes.c_char_p, []),
    ("bclx_version", ctypes.c_char_p, []),
    ("bclx_create", _I32, [_I32, _P, ctypes.POINTER(_P)]),
    ("bclx_destroy", _I32, [_P]),
    ("bclx_set_stream", _I32, [_P, _P]),
    ("bclx_synchronize", _I32, [_P]),
    ("bclx_set_allocator", _I32, [_P, _P, _P, _P]),
    ("bclx_spgemm", _I32, [_P, ctypes.POINTER(bclx_csr), ctypes.POINTER(bclx_csr), ctypes.POINTER(bclx_csr)]),
    ("bclx_spgemm_noalloc", _I32, [_P, ctypes.POINTER(bclx_csr), ctypes.POINTER(bclx_csr), ctypes.POINTER(bclx_csr)]),
    ("bclx_spgemm_symbolic", _I32, [_P, ctypes.POINTER(bclx_csr), ctypes.POINTER(bclx_csr), ctypes.POINTER(_P),
                                    ctypes.POINTER(_I64)]),
    ("bclx_spgemm_numeric", _I32, [_P, ctypes.POINTER(bclx_csr)]),
    ("bclx_spgemm_plan_destroy", _I32, [_P]),
    ("bclx_spgemm_plan_rows", _I32, [_P, _P, _P]),
    ("bclx_accumulate_tiles", _I32, [_P, _I32, ctypes.POINTER(bclx_csr), ctypes.POINTER(bclx_csr)]),
    ("bclx_stage_plan_create", _I32, [_P, _I32, ctypes.POINTER(bclx_csr), ctypes.POINTER(bclx_csr),
                                      ctypes.POINTER(bclx_stage_opts), ctypes.POINTER(_P)]),
    ("bclx_stage_execute", _I32, [_P]),
    ("bclx_stage_result", _I32, [_P, _I32, ctypes.POINTER(bclx_csr)]),
    ("bclx_stage_stats", _I32, [_P, _P, _I32]),
    ("bclx_stage_predictions", _I32, [_P, _I64, _P, _P, _P, _P, _P, ctypes.POINTER(_I64)]),
    ("bclx_stage_plan_destroy", _I32, [_P]),
    ("bclx_row_features", _I32, [_P, ctypes.POINTER(bclx_csr), ctypes.POINTER(bclx_csr), _P, _P]),
    ("bclx_mlp_forward", _I32, [_P, _I64, _P, _P, _P]),
    ("bclx_mlp_grad", _I32, [_P, _I64, _P, _P, _P, _P, _P]),
    ("bclx_mlp_adam", _I32, [_P, _P, _P, _P, _P, _I32, _F32, _F32, _F32, _F32]),
    ("bclx_spmm", _I32, [_P, ctypes.POINTER(bclx_csr), _P, _I32, _I64, _P, _I64, _I64, _F32, _F32]),
    ("bclx_gemm", _I32, [_P, _I32, _I32, _I64, _I64, _I64, _F32, _P, _I64, _P, _I64, _F32, _P, _I64, _I32]),
    ("bclx_spmm_f64", _I32, [_P, ctypes.POINTER(bclx_csr), _P, _I64, _P, _I64, _I64, ctypes.c_double, ctypes.c_double]),
    ("bclx_gemm_f64", _I32, [_P, _I32, _I32, _I64, _I64, _I64, ctypes.c_double, _P, _I64, _P, _I64, ctypes.c_double, _P,
                             _I64]),
    ("bclx_gemm_f64_ref", _I32, [_P, _I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64]),
    ("bclx_csr_split_grid", _I32, [_P, ctypes.POINTER(bclx_csr), _I64, _I64, _I32, ctypes.POINTER(bclx_csr)]),
    ("bclx_dense_binary_op", _I32, [_P, _I32, _I32, _I64, _I64, _P, _I64, _P, _I64, _P, _I64]),
    ("bclx_mm_info", _I32, [ctypes.c_char_p, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    ("bclx_mm_read", _I32, [ctypes.c_char_p, _I32, _I32, _I32, ctypes.POINTER(bclx_csr)]),
    ("bclx_mm_write", _I32, [ctypes.c_char_p, ctypes.POINTER(bclx_csr)]),
    ("bclx_host_free", None, [_P]),
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (python -m bcl_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise BclxError(status, lib().bclx_last_error().decode())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
