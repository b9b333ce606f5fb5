"""bclx_spgemm_noalloc (one call, no device allocation, workspace in the library's fixed scratch array) against
bclx_spgemm_symbolic + bclx_spgemm_numeric on the same inputs: identical row_ptr, columns and value bits, on two
calls in a row (the second on a smaller input, so it runs over what the first left in the scratch), for fp32 / int32
and fp64 / int64 inputs. Also: a C too small fails with BCLX_ERR_OVERFLOW and reports the entries needed, and an A with
more rows than the scratch holds is refused before any work."""
import ctypes

import numpy as np
import pytest
import torch

from bcl_amd import gen
from bcl_amd._lib import BCLX_F32, BCLX_F64, BclxError, bclx_csr, call
from bcl_amd.csr import CSR
from bcl_amd.device import DeviceCSR, spgemm

pytestmark = pytest.mark.gpu
BCLX_ERR_OVERFLOW = 4


def noalloc(handle, A: DeviceCSR, B: DeviceCSR, cap: int, ptr_dtype) -> DeviceCSR:
    dev = A.vals.device
    C = DeviceCSR(A.m, B.n, torch.full((A.m + 1,), -7, dtype=ptr_dtype, device=dev),
                  torch.full((max(cap, 1),), -7, dtype=torch.int32, device=dev),
                  torch.full((max(cap, 1),), -7.0, dtype=A.vals.dtype, device=dev))
    d = C.desc()
    d.nnz = cap
    da, db = A.desc(), B.desc()
    call("bclx_spgemm_noalloc", handle.h, ctypes.byref(da), ctypes.byref(db), ctypes.byref(d))
    assert d.m == A.m and d.n == B.n
    C.col_idx, C.vals = C.col_idx[:d.nnz], C.vals[:d.nnz]
    return C


def same(x: DeviceCSR, y: DeviceCSR):
    assert x.nnz == y.nnz
    assert torch.equal(x.row_ptr.long(), y.row_ptr.long())
    assert torch.equal(x.col_idx, y.col_idx)
    assert torch.equal(x.vals.view(torch.uint8), y.vals.view(torch.uint8))


@pytest.mark.parametrize("f64,ptr64", [(False, False), (True, True)])
def test_noalloc_matches_two_phase(handle, f64, ptr64):
    pd = torch.int64 if ptr64 else torch.int32
    first = gen.rmat(15, seed=3)   # hub rows: heavy and sorted bins
    second = gen.uniform(3000, 3000, 8, seed=4)  # smaller, runs over the first call's leftovers
    for a in (first, second):
        if f64:
            a = CSR(a.m, a.n, a.row_ptr, a.col_idx, a.vals.astype(np.float64))
        print(f"[noalloc] {a.m}x{a.n} nnz={a.nnz} f64={f64} ptr64={ptr64}", flush=True)
        A = DeviceCSR.from_host(a, ptr_dtype=pd)
        want = spgemm(handle, A, A, ptr_dtype=pd)
        got = noalloc(handle, A, A, want.nnz + 100, pd)
        torch.cuda.synchronize()
        same(got, want)


def test_noalloc_reports_short_output_and_large_inputs(handle):
    a = gen.uniform(2000, 2000, 8, seed=5)
    A = DeviceCSR.from_host(a)
    want = spgemm(handle, A, A, ptr_dtype=torch.int32)
    d = bclx_csr(A.m, A.n, 10, 4, BCLX_F32, torch.zeros(A.m + 1, dtype=torch.int32, device="cuda").data_ptr(),
                 torch.zeros(10, dtype=torch.int32, device="cuda").data_ptr(),
                 torch.zeros(10, dtype=torch.float32, device="cuda").data_ptr())
    da = A.desc()
    with pytest.raises(BclxError) as e:
        call("bclx_spgemm_noalloc", handle.h, ctypes.byref(da), ctypes.byref(da), ctypes.byref(d))
    assert e.value.status == BCLX_ERR_OVERFLOW and d.nnz == want.nnz
    big = CSR.empty(1 << 21, 16)
    Bg = DeviceCSR.from_host(CSR.empty(16, 16))
    Ag = DeviceCSR.from_host(big)
    dg = bclx_csr(big.m, 16, 0, 4, BCLX_F32, torch.zeros(big.m + 1, dtype=torch.int32, device="cuda").data_ptr(), 0, 0)
    dag, dbg = Ag.desc(), Bg.desc()
    with pytest.raises(BclxError) as e:
        call("bclx_spgemm_noalloc", handle.h, ctypes.byref(dag), ctypes.byref(dbg), ctypes.byref(dg))
    assert e.value.status == BCLX_ERR_OVERFLOW and "scratch" in str(e.value)
    # the scratch is still usable after the refusals
    same(noalloc(handle, A, A, want.nnz, torch.int32), want)
