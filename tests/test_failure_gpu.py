"""Failure handling, forced from the host (no kernel is made to fault or hang):

- a stage plan executed after its inputs' sparsity structure changed: a structure that outgrows what the plan was
  sized for must fail with BCLX_ERR_OVERFLOW (never hand out a partial result), and one that still fits must give
  the oracle's result for the new inputs;
- an int32 row_ptr overflow: a product with more than 2^31 - 1 entries (a dense outer product counted by the
  symbolic pass) requested with a 4-byte row_ptr fails with BCLX_ERR_OVERFLOW before anything is written;
- allocation failure: a caller-supplied allocator that returns NULL makes the call fail with BCLX_ERR_ALLOC, and a
  working one is the one the library allocates through (plans and the outputs of bclx_spgemm).
The reference throws std::runtime_error on these (SPMatrix.hpp:251-253, gemm.hpp:123-125).
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle_py as O
from bcl_amd import predictor
from bcl_amd._lib import BCLX_F32, BclxError, bclx_csr, call
from bcl_amd.csr import CSR
from bcl_amd.device import DeviceCSR, Handle, spgemm
from bcl_amd.stage import BATCHED_EXACT, LEARNED, PLAIN, StagePlan, d2h
from conftest import spgemm_value_bound

pytestmark = pytest.mark.gpu

BCLX_ERR_ALLOC, BCLX_ERR_OVERFLOW = 3, 4


def _abs(a: CSR) -> CSR:
    return CSR(a.m, a.n, a.row_ptr, a.col_idx, np.abs(a.vals))


def _check(c: CSR, a: CSR, b: CSR):
    ref = O.spgemm(a, b)
    assert c.nnz == ref.nnz
    np.testing.assert_array_equal(c.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(c.col_idx, ref.col_idx)
    c64 = O.spgemm_f64(a, b)
    absum = O.spgemm_f64(_abs(a), _abs(b))
    err = np.abs(c.vals.astype(np.float64) - c64.vals)
    assert (err <= spgemm_value_bound(a, b, c64.vals, absum.vals, np.float32)).all()


def _structure_inputs(seed=5):
    """A (4096 x 4096) has 8 entries per row, at columns 0..7. B's rows 0..7 hold one entry each, rows 8..15 512
    entries each, rows 16..23 one entry each at column 5, and the other rows two entries. Adding 8 to A's columns
    makes every C row the union of B's long rows (about 3.2 K entries instead of 8); adding 16 makes it one entry."""
    rng = np.random.default_rng(seed)
    m = n = 4096
    a = CSR(m, n, np.arange(0, 8 * m + 1, 8, dtype=np.int64), np.tile(np.arange(8, dtype=np.int32), m),
            rng.uniform(-1, 1, 8 * m).astype(np.float32))
    rows = []
    for r in range(n):
        if r < 8:
            cols = [8 * r]
        elif r < 16:
            cols = np.sort(rng.choice(n, 512, replace=False))
        elif r < 24:
            cols = [5]
        else:
            cols = np.sort(rng.choice(n, 2, replace=False))
        rows.append(np.asarray(cols, np.int32))
    rp = np.concatenate([[0], np.cumsum([len(x) for x in rows])]).astype(np.int64)
    b = CSR(n, n, rp, np.concatenate(rows), rng.uniform(-1, 1, int(rp[-1])).astype(np.float32))
    return a, b


def _shifted(a: CSR, by: int) -> CSR:
    return CSR(a.m, a.n, a.row_ptr, (a.col_idx + by).astype(np.int32), a.vals)


@pytest.mark.parametrize("mode", [LEARNED, BATCHED_EXACT, PLAIN])
def test_stage_after_structure_change(handle, mode):
    a, b = _structure_inputs()
    A, B = DeviceCSR.from_host(a), DeviceCSR.from_host(b)
    print(f"[failure] structure change, mode {mode}", flush=True)
    plan = StagePlan(handle, [A], [B], mode=mode, params=predictor.load())
    plan.execute()
    torch.cuda.synchronize()
    _check(plan.result_host(0), a, b)
    # grown: more output than the plan holds -> an error, never a partial result
    A.col_idx.add_(8)
    torch.cuda.synchronize()
    with pytest.raises(BclxError) as e:
        plan.execute()
        torch.cuda.synchronize()
        plan.result_host(0)
    assert e.value.status == BCLX_ERR_OVERFLOW, str(e.value)
    # shrunk (one entry per row): the plan still fits and recomputes the new product
    A.col_idx.add_(8)
    torch.cuda.synchronize()
    plan.execute()
    torch.cuda.synchronize()
    _check(plan.result_host(0), _shifted(a, 16), b)
    # back to the original inputs: the failed execution left nothing behind
    A.col_idx.sub_(16)
    torch.cuda.synchronize()
    plan.execute()
    torch.cuda.synchronize()
    _check(plan.result_host(0), a, b)
    plan.close()


def test_int32_row_ptr_overflow(handle):
    """C = x * y^T with x, y all-ones vectors of 2^16 entries: 2^32 entries, counted exactly by the symbolic pass."""
    m = 1 << 16
    x = CSR(m, 1, np.arange(m + 1, dtype=np.int64), np.zeros(m, np.int32), np.ones(m, np.float32))
    y = CSR(1, m, np.array([0, m], np.int64), np.arange(m, dtype=np.int32), np.ones(m, np.float32))
    X, Y = DeviceCSR.from_host(x), DeviceCSR.from_host(y)
    dx, dy = X.desc(), Y.desc()
    plan = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    call("bclx_spgemm_symbolic", handle.h, ctypes.byref(dx), ctypes.byref(dy), ctypes.byref(plan), ctypes.byref(nnz))
    try:
        assert nnz.value == 1 << 32
        rp = torch.zeros(m + 1, dtype=torch.int32, device="cuda")
        dc = bclx_csr(m, m, nnz.value, 4, BCLX_F32, rp.data_ptr(), 0, 0)
        with pytest.raises(BclxError) as e:
            call("bclx_spgemm_numeric", plan, ctypes.byref(dc))
        assert e.value.status == BCLX_ERR_OVERFLOW
        torch.cuda.synchronize()
        assert int(rp.abs().sum()) == 0, "nothing may be written when the request is refused"
    finally:
        call("bclx_spgemm_plan_destroy", plan)
    out = bclx_csr()
    out.ptr_width = 4
    with pytest.raises(BclxError) as e:
        call("bclx_spgemm", handle.h, ctypes.byref(dx), ctypes.byref(dy), ctypes.byref(out))
    assert e.value.status == BCLX_ERR_OVERFLOW


class _HipAllocator:
    """A caller-supplied device allocator (hipMalloc / hipFree underneath) that counts its calls and can be told to
    refuse."""
    ALLOC = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)
    FREE = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMalloc.restype = ctypes.c_int
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipFree.restype = ctypes.c_int
        self.allocs = self.frees = 0
        self.refuse = False
        self.live = set()
        self._a = self.ALLOC(self._alloc)
        self._f = self.FREE(self._free)

    def _alloc(self, nbytes, user):
        self.allocs += 1
        if self.refuse:
            return None
        p = ctypes.c_void_p()
        if self.hip.hipMalloc(ctypes.byref(p), nbytes) != 0:
            return None
        self.live.add(p.value)
        return p.value

    def _free(self, ptr, user):
        self.frees += 1
        self.live.discard(ptr)
        self.hip.hipFree(ptr)

    def install(self, h):
        call("bclx_set_allocator", h.h, ctypes.cast(self._a, ctypes.c_void_p), ctypes.cast(self._f, ctypes.c_void_p),
             None)


def test_allocator_callback_and_allocation_failure():
    from bcl_amd import gen
    a = gen.rmat(12, seed=9)
    A = DeviceCSR.from_host(a)
    h = Handle(0)
    al = _HipAllocator()
    try:
        al.install(h)
        # refusing allocator: every entry point that needs memory fails cleanly with BCLX_ERR_ALLOC
        al.refuse = True
        with pytest.raises(BclxError) as e:
            spgemm(h, A, A)
        assert e.value.status == BCLX_ERR_ALLOC, str(e.value)
        with pytest.raises(BclxError) as e:
            StagePlan(h, [A], [A], mode=LEARNED, params=predictor.load())
        assert e.value.status == BCLX_ERR_ALLOC, str(e.value)
        # working allocator: the library's buffers and bclx_spgemm's outputs come from it
        al.refuse = False
        before = al.allocs
        da = A.desc()
        out = bclx_csr()
        out.ptr_width = 8
        call("bclx_spgemm", h.h, ctypes.byref(da), ctypes.byref(da), ctypes.byref(out))
        assert al.allocs - before >= 4, "plan workspace and the three output arrays go through the allocator"
        assert {out.row_ptr, out.col_idx, out.vals} <= al.live, "the outputs are the caller's, still allocated"
        rp = np.zeros(out.m + 1, np.int64)
        col = np.zeros(out.nnz, np.int32)
        val = np.zeros(out.nnz, np.float32)
        d2h(rp, out.row_ptr)
        d2h(col, out.col_idx)
        d2h(val, out.vals)
        _check(CSR(out.m, out.n, rp, col, val), a, a)
        for ptr in (out.row_ptr, out.col_idx, out.vals):
            al._free(ptr, None)
        plan = StagePlan(h, [A], [A], mode=LEARNED, params=predictor.load())
        plan.execute()
        torch.cuda.synchronize()
        _check(plan.result_host(0), a, a)
        plan.close()
    finally:
        call("bclx_set_allocator", h.h, None, None, None)
        h.close()
    assert al.allocs > 0 and not al.live, "every buffer made through the allocator was given back through it"
