"""ctypes bindings of the CPU oracle (liboracle.so) and of the reference build (_ref/libbclref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py,
always as the checker or the timed CPU baseline, never as the product path.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))
from bcl_amd.csr import CSR  # noqa: E402

_P = ctypes.c_void_p
_I64 = ctypes.c_int64


class _OCSR(ctypes.Structure):
    _fields_ = [("m", _I64), ("n", _I64), ("nnz", _I64), ("row_ptr", _P), ("col_idx", _P), ("vals", _P)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        sp = [_I64, _I64, _I64, _P, _P, _P, _P, _P, _P]
        L.oracle_spgemm_f32.argtypes = sp + [_P, ctypes.c_int, _P]
        L.oracle_spgemm_f64.argtypes = sp + [_P, ctypes.c_int, _P]
        L.oracle_spgemm_f32_to_f64.argtypes = sp + [_P]
        L.oracle_spgemm_rowhash_f32.argtypes = sp + [ctypes.c_int, _P]
        L.oracle_spgemm_rowhash_f64.argtypes = sp + [ctypes.c_int, _P]
        L.oracle_spgemm_row_nnz.argtypes = [_I64, _I64, _I64, _P, _P, _P, _P, ctypes.c_int, _P, _P]
        L.oracle_csr_free.argtypes = [_P]
        L.oracle_spmm_f32.argtypes = [_I64, _I64, _I64, _P, _P, _P, _P, _I64, _P, _I64]
        L.oracle_spmm_f32.restype = None
        L.oracle_spmm_f64.argtypes = [_I64, _I64, _I64, _P, _P, _P, _P, _P, _I64, _P, _I64, _P, ctypes.c_int]
        L.oracle_spmm_f64.restype = None
        L.oracle_gemm_f64.argtypes = [_I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64, _I64, _I64, ctypes.c_int]
        L.oracle_gemm_f64.restype = None
        L.oracle_accumulate_f32.argtypes = [_I64, _I64, ctypes.c_int, _P, _P, _P, ctypes.c_int, _P]
        L.oracle_mlp_forward_backward.argtypes = [_I64, _P, _P, _P, _P, _P, _P]
        L.oracle_mlp_forward_backward.restype = None
        _lib = L
    return _lib


def have_ref() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libbclref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libbclref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (needs the reference source tree): `make -C oracle ref`")
        R = ctypes.CDLL(path)
        sp = [_I64, _I64, _I64, _I64, _P, _P, _P, _I64, _P, _P, _P, _I64, _P, _P, _P]
        R.ref_spgemm_f32.argtypes = sp
        R.ref_spgemm_f32.restype = _P
        R.ref_spgemm_f64.argtypes = sp
        R.ref_spgemm_f64.restype = _P
        R.ref_accumulate_f32.argtypes = [_I64, _I64, ctypes.c_int, _P, _P, _P, _P]
        R.ref_accumulate_f32.restype = _P
        R.ref_result_nnz.argtypes = [_P, ctypes.c_int]
        R.ref_result_nnz.restype = _I64
        R.ref_result_copy_free.argtypes = [_P, ctypes.c_int, _P, _P, _P]
        R.ref_result_copy_free.restype = None
        R.ref_spmm_f32.argtypes = [_I64, _I64, _I64, _I64, _P, _P, _P, _P, _P, _I64, _I64]
        R.ref_binary_roundtrip.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        R.ref_binary_roundtrip.restype = _I64
        R.ref_spmm_f32.restype = None
        for nm in ("ref_mm_read_f32", "ref_mm_read_f64"):
            getattr(R, nm).argtypes = [ctypes.c_char_p, ctypes.c_int]
            getattr(R, nm).restype = _P
        R.ref_error.argtypes = []
        R.ref_error.restype = ctypes.c_char_p
        R.ref_result_m.argtypes = [_P, ctypes.c_int]
        R.ref_result_m.restype = _I64
        R.ref_result_n.argtypes = [_P, ctypes.c_int]
        R.ref_result_n.restype = _I64
        R.ref_mm_info.argtypes = [ctypes.c_char_p, _P]
        R.ref_mm_info.restype = None
        R.ref_mm_write_f32.argtypes = [ctypes.c_char_p, _I64, _I64, _I64, _P, _P, _P]
        R.ref_mm_write_f32.restype = None
        R.ref_mmap_slice_f32.argtypes = [ctypes.c_char_p, _I64, _I64, _I64, _I64]
        R.ref_mmap_slice_f32.restype = _P
        R.ref_slice_f32.argtypes = [_I64, _I64, _I64, _P, _P, _P, _I64, _I64, _I64, _I64]
        R.ref_slice_f32.restype = _P
        _ref = R
    return _ref


def _take(o: _OCSR, dtype) -> CSR:
    m, n, nnz = o.m, o.n, o.nnz
    rp = np.ctypeslib.as_array(ctypes.cast(o.row_ptr, ctypes.POINTER(ctypes.c_int64)), shape=(m + 1,)).copy()
    if nnz:
        col = np.ctypeslib.as_array(ctypes.cast(o.col_idx, ctypes.POINTER(ctypes.c_int32)), shape=(nnz,)).copy()
        ct = ctypes.c_float if dtype == np.float32 else ctypes.c_double
        val = np.ctypeslib.as_array(ctypes.cast(o.vals, ctypes.POINTER(ct)), shape=(nnz,)).copy()
    else:
        col = np.zeros(0, np.int32)
        val = np.zeros(0, dtype)
    lib().oracle_csr_free(ctypes.byref(o))
    return CSR(m, n, rp, col, val)


def _c64(a: CSR):
    return (np.ascontiguousarray(a.row_ptr, np.int64), np.ascontiguousarray(a.col_idx, np.int32))


def spgemm(a: CSR, b: CSR, c0: CSR | None = None, dead_row_alloc: bool = False) -> CSR:
    """BCL::gemm(CSR, CSR, CSR) restated: c0 + a*b in the dtype of a (float32 or float64)."""
    dt = a.vals.dtype
    arp, acol = _c64(a)
    brp, bcol = _c64(b)
    av = np.ascontiguousarray(a.vals)
    bv = np.ascontiguousarray(b.vals.astype(dt))
    keep = []
    c0p = None
    if c0 is not None:
        crp, ccol = _c64(c0)
        cv = np.ascontiguousarray(c0.vals.astype(dt))
        keep = [crp, ccol, cv]
        c0s = _OCSR(c0.m, c0.n, c0.nnz, _ptr(crp), _ptr(ccol), _ptr(cv))
        c0p = ctypes.byref(c0s)
    o = _OCSR()
    fn = lib().oracle_spgemm_f32 if dt == np.float32 else lib().oracle_spgemm_f64
    fn(a.m, a.n, b.n, _ptr(arp), _ptr(acol), _ptr(av), _ptr(brp), _ptr(bcol), _ptr(bv), c0p, int(dead_row_alloc),
       ctypes.byref(o))
    del keep
    return _take(o, dt)


def spgemm_f64(a: CSR, b: CSR) -> CSR:
    """FP64 value oracle: the same algorithm accumulating in float64 (fp32 inputs widened exactly)."""
    arp, acol = _c64(a)
    brp, bcol = _c64(b)
    o = _OCSR()
    if a.vals.dtype == np.float32:
        av = np.ascontiguousarray(a.vals)
        bv = np.ascontiguousarray(b.vals, np.float32)
        lib().oracle_spgemm_f32_to_f64(a.m, a.n, b.n, _ptr(arp), _ptr(acol), _ptr(av), _ptr(brp), _ptr(bcol),
                                       _ptr(bv), ctypes.byref(o))
    else:
        av = np.ascontiguousarray(a.vals, np.float64)
        bv = np.ascontiguousarray(b.vals, np.float64)
        lib().oracle_spgemm_f64(a.m, a.n, b.n, _ptr(arp), _ptr(acol), _ptr(av), _ptr(brp), _ptr(bcol), _ptr(bv),
                                None, 0, ctypes.byref(o))
    return _take(o, np.float64)


def spgemm_rowhash(a: CSR, b: CSR, nthreads: int = 0) -> CSR:
    dt = a.vals.dtype
    arp, acol = _c64(a)
    brp, bcol = _c64(b)
    av = np.ascontiguousarray(a.vals)
    bv = np.ascontiguousarray(b.vals.astype(dt))
    o = _OCSR()
    fn = lib().oracle_spgemm_rowhash_f32 if dt == np.float32 else lib().oracle_spgemm_rowhash_f64
    fn(a.m, a.n, b.n, _ptr(arp), _ptr(acol), _ptr(av), _ptr(brp), _ptr(bcol), _ptr(bv), nthreads or os.cpu_count(),
       ctypes.byref(o))
    return _take(o, dt)


def row_nnz(a: CSR, b: CSR, nthreads: int = 0):
    """(nnz(C_i), products_i) per row, both int64."""
    arp, acol = _c64(a)
    brp, bcol = _c64(b)
    out = np.zeros(a.m, np.int64)
    prod = np.zeros(a.m, np.int64)
    lib().oracle_spgemm_row_nnz(a.m, a.n, b.n, _ptr(arp), _ptr(acol), _ptr(brp), _ptr(bcol),
                                nthreads or os.cpu_count(), _ptr(out), _ptr(prod))
    return out, prod


def spmm(a: CSR, b: np.ndarray, c: np.ndarray) -> None:
    """In place c += a @ b (fp32, reference loop order)."""
    arp, acol = _c64(a)
    lib().oracle_spmm_f32(a.m, b.shape[1], a.n, _ptr(arp), _ptr(acol), _ptr(a.vals), _ptr(b), b.shape[1], _ptr(c),
                          c.shape[1])


def spmm_f64(a: CSR, b: np.ndarray, c0: np.ndarray | None = None, nthreads: int = 0) -> np.ndarray:
    """float64 c0 + a @ b; b is float32 or bf16 bits (uint16); c0 (optional) is taken as float32."""
    arp, acol = _c64(a)
    assert a.vals.dtype == np.float32 and b.dtype in (np.float32, np.uint16), "fp32 A and fp32 / bf16-bits B"
    if c0 is not None:
        c0 = np.ascontiguousarray(c0, np.float32)
        assert c0.shape == (a.m, b.shape[1])
    n = b.shape[1]
    out = np.zeros((a.m, n), np.float64)
    bf = b if b.dtype == np.float32 else None
    bh = b if b.dtype == np.uint16 else None
    lib().oracle_spmm_f64(a.m, n, a.n, _ptr(arp), _ptr(acol), _ptr(a.vals), _ptr(bf), _ptr(bh), n,
                          _ptr(c0), n, _ptr(out), nthreads or os.cpu_count())
    return out


def gemm_f64(a: np.ndarray, b: np.ndarray, c0: np.ndarray | None = None, rows=None, nthreads: int = 0) -> np.ndarray:
    """float64 (c0 +) a @ b for the row range rows=(r0, r1) (all rows by default); returns those rows."""
    m, k = a.shape
    n = b.shape[1]
    r0, r1 = rows if rows is not None else (0, m)
    out = np.zeros((m, n), np.float64) if rows is None else np.zeros((m, n), np.float64)
    if c0 is not None:
        out[r0:r1] = c0[r0:r1]
    lib().oracle_gemm_f64(m, n, k, _ptr(a), a.shape[1], _ptr(b), n, _ptr(out), n, r0, r1, nthreads or os.cpu_count())
    return out[r0:r1]


def accumulate(tiles, m: int, n: int, nthreads: int = 0) -> CSR:
    """SparseHashAccumulator::get_matrix over fp32 tiles (all m x n, offset 0)."""
    rps = [np.ascontiguousarray(t.row_ptr, np.int64) for t in tiles]
    cols = [np.ascontiguousarray(t.col_idx, np.int32) for t in tiles]
    vals = [np.ascontiguousarray(t.vals, np.float32) for t in tiles]
    arr = lambda xs: (_P * len(xs))(*[x.ctypes.data for x in xs])
    o = _OCSR()
    lib().oracle_accumulate_f32(m, n, len(tiles), arr(rps), arr(cols), arr(vals), nthreads or os.cpu_count(),
                                ctypes.byref(o))
    return _take(o, np.float32)


def mlp(x: np.ndarray, params: np.ndarray, target: np.ndarray | None = None):
    """float64 forward (and gradient of 0.5*mean((y - t)^2) when target is given). Returns (y, grads, loss)."""
    x = np.ascontiguousarray(x, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    y = np.zeros(x.shape[0], np.float64)
    grads = np.zeros(5313, np.float64) if target is not None else None
    loss = ctypes.c_double(0)
    t = np.ascontiguousarray(target, np.float32) if target is not None else None
    lib().oracle_mlp_forward_backward(x.shape[0], _ptr(x), _ptr(t), _ptr(params), _ptr(y), _ptr(grads),
                                      ctypes.byref(loss))
    return y, grads, loss.value


# ---- the reference itself (oracle/_ref, built from the reference headers) ------------------------------------

def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _ref_take(r, m, n, f64: bool) -> CSR:
    R = ref()
    nnz = R.ref_result_nnz(r, int(f64))
    rp = np.zeros(m + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float64 if f64 else np.float32)
    R.ref_result_copy_free(r, int(f64), _ptr(rp), _ptr(col), _ptr(val))
    return CSR(m, n, rp.astype(np.int64), col[:nnz], val[:nnz])


def ref_spgemm(a: CSR, b: CSR, c0: CSR | None = None) -> CSR:
    """The reference BCL::gemm(CSRMatrix, CSRMatrix, CSRMatrix) (c0 += a * b) in the dtype of a."""
    f64 = a.vals.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    arp, acol, av = _i32(a.row_ptr), _i32(a.col_idx), np.ascontiguousarray(a.vals, dt)
    brp, bcol, bv = _i32(b.row_ptr), _i32(b.col_idx), np.ascontiguousarray(b.vals, dt)
    if c0 is not None:
        crp, ccol, cv, cn = _i32(c0.row_ptr), _i32(c0.col_idx), np.ascontiguousarray(c0.vals, dt), c0.nnz
    else:
        crp = ccol = cv = None
        cn = 0
    fn = ref().ref_spgemm_f64 if f64 else ref().ref_spgemm_f32
    r = fn(a.m, a.n, b.n, a.nnz, _ptr(arp), _ptr(acol), _ptr(av), b.nnz, _ptr(brp), _ptr(bcol), _ptr(bv), cn,
           _ptr(crp), _ptr(ccol), _ptr(cv))
    return _ref_take(r, a.m, b.n, f64)


def ref_accumulate(tiles, m: int, n: int) -> CSR:
    rps = [_i32(t.row_ptr) for t in tiles]
    cols = [_i32(t.col_idx) if t.nnz else np.zeros(1, np.int32) for t in tiles]
    vals = [np.ascontiguousarray(t.vals, np.float32) if t.nnz else np.zeros(1, np.float32) for t in tiles]
    nnzs = np.array([t.nnz for t in tiles], np.int64)
    arr = lambda xs: (_P * len(xs))(*[x.ctypes.data for x in xs])
    r = ref().ref_accumulate_f32(m, n, len(tiles), _ptr(nnzs), arr(rps), arr(cols), arr(vals))
    return _ref_take(r, m, n, False)


def ref_spmm(a: CSR, b: np.ndarray, c: np.ndarray) -> None:
    """The reference spmm_wrapper: c += a @ b in place (fp32)."""
    arp, acol = _i32(a.row_ptr), _i32(a.col_idx)
    av = np.ascontiguousarray(a.vals, np.float32)
    ref().ref_spmm_f32(a.m, b.shape[1], a.n, a.nnz, _ptr(c), _ptr(b), _ptr(av), _ptr(arp), _ptr(acol), b.shape[1],
                       c.shape[1])


class RefError(RuntimeError):
    """An exception the reference threw (its message)."""


def ref_mm_read(path: str, one_indexed: bool = True, dtype=np.float32) -> CSR:
    """The reference CSRMatrix<T, int>::read_MatrixMarket (over matrix_io::mmread); RefError on its exceptions."""
    f64 = dtype == np.float64
    R = ref()
    r = (R.ref_mm_read_f64 if f64 else R.ref_mm_read_f32)(path.encode(), int(one_indexed))
    if not r:
        raise RefError(R.ref_error().decode())
    return _ref_take(r, R.ref_result_m(r, int(f64)), R.ref_result_n(r, int(f64)), f64)


def ref_mm_info(path: str):
    out = np.zeros(3, np.int64)
    ref().ref_mm_info(path.encode(), _ptr(out))
    return tuple(int(x) for x in out)


def ref_mm_write(path: str, a: CSR) -> None:
    ref().ref_mm_write_f32(path.encode(), a.m, a.n, a.nnz, _ptr(_i32(a.row_ptr)), _ptr(_i32(a.col_idx)),
                           _ptr(np.ascontiguousarray(a.vals, np.float32)))


def ref_mmap_slice(path: str, imin: int, imax: int, jmin: int, jmax: int) -> CSR:
    """The reference CSRMatrixMemoryMapped<float, int>(path).get_slice_impl_(imin, imax, jmin, jmax)."""
    r = ref().ref_mmap_slice_f32(path.encode(), imin, imax, jmin, jmax)
    return _ref_take(r, ref().ref_result_m(r, 0), ref().ref_result_n(r, 0), False)


def ref_slice(a: CSR, imin: int, imax: int, jmin: int, jmax: int) -> CSR:
    """The reference CSRMatrix<float, int>::get_slice_impl_ of an in-memory matrix."""
    r = ref().ref_slice_f32(a.m, a.n, a.nnz, _ptr(_i32(a.row_ptr)), _ptr(_i32(a.col_idx)),
                            _ptr(np.ascontiguousarray(a.vals, np.float32)), imin, imax, jmin, jmax)
    return _ref_take(r, ref().ref_result_m(r, 0), ref().ref_result_n(r, 0), False)
