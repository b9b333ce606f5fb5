"""GPU parity of SpMM (bclx_spmm, BCL's spmm_wrapper: C += A * B) against the oracle.

Accumulation is fp32, so values are compared with the float64 oracle through a magnitude bound:
|c - c64| <= 1e-5 * (|beta c0| + |alpha| sum_k |a_ik b_kj|)  (the stated SpMM tolerance, about 170x the fp32 unit
roundoff; row lengths here stay below a few thousand terms). For bf16 B the float64 oracle uses the same bf16
values, and the bf16 rounding error against the fp32-B result is reported separately.
"""
import os

import numpy as np
import pytest
import torch

import oracle_py as O
from bcl_amd import gen
from bcl_amd.csr import CSR
from bcl_amd.device import DeviceCSR, spmm

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _abs(a):
    return CSR(a.m, a.n, a.row_ptr, a.col_idx, np.abs(a.vals))


def check(handle, a, b, c0, n=None, alpha=1.0, beta=1.0, bf16=False, ptr_dtype=torch.int32):
    n = b.shape[1] if n is None else n
    print(f"[spmm] A {a.m}x{a.n} nnz={a.nnz} n={n} ld={b.shape[1]}/{c0.shape[1]} bf16={bf16}", flush=True)
    A = DeviceCSR.from_host(a, ptr_dtype=ptr_dtype)
    Bt = torch.from_numpy(b).cuda()
    if bf16:
        Bt = Bt.to(torch.bfloat16)
        bh = Bt.view(torch.int16).cpu().numpy().view(np.uint16)
    Ct = torch.from_numpy(c0.copy()).cuda()
    spmm(handle, A, Bt, Ct, n=n, alpha=alpha, beta=beta)
    torch.cuda.synchronize()
    c = Ct.cpu().numpy()
    bsrc = bh if bf16 else np.ascontiguousarray(b)
    prod64 = O.spmm_f64(a, bsrc)[:, :n] if bsrc.shape[1] == n else O.spmm_f64(a, np.ascontiguousarray(bsrc[:, :n]))
    expect = beta * c0[:, :n].astype(np.float64) + alpha * prod64
    absb = np.abs(b[:, :n]).astype(np.float32)
    mag = abs(beta) * np.abs(c0[:, :n]) + abs(alpha) * O.spmm_f64(_abs(a), np.ascontiguousarray(absb))
    err = np.abs(c[:, :n] - expect)
    assert np.all(err <= 1e-5 * mag + 1e-30), float((err / np.maximum(mag, 1e-30)).max())
    np.testing.assert_array_equal(c[:, n:], c0[:, n:])  # padding columns untouched
    return c


def test_golden_reference_case(handle):
    d = np.load(os.path.join(GOLDEN, "spmm_rmat9.npz"), allow_pickle=False)
    m, k = (int(x) for x in d["a_shape"])
    a = CSR(m, k, d["a_row_ptr"], d["a_col_idx"], d["a_vals"])
    n = int(d["n"][0])
    c = check(handle, a, d["b"], d["c0"], n=n)
    ref = d["c"]
    assert np.all(np.abs(c - ref) <= 1e-5 * (np.abs(ref) + 1))


@pytest.mark.parametrize("bf16", [False, True])
def test_rmat_128(handle, bf16):
    a = gen.rmat(14, seed=3)
    rng = np.random.default_rng(1)
    b = rng.standard_normal((a.n, 128)).astype(np.float32)
    c0 = np.zeros((a.m, 128), np.float32)
    c = check(handle, a, b, c0, bf16=bf16)
    if bf16:
        exact = O.spmm_f64(a, b)
        rel = np.abs(c - exact).max() / np.abs(exact).max()
        print(f"[spmm] bf16-B max error vs fp32-B FP64 result, relative to max |C|: {rel:.3e}")
        assert rel < 1e-2


@pytest.mark.parametrize("bf16", [False, True])
def test_alpha_beta_int64_and_odd_widths(handle, bf16):
    """Widths that take the paired-entry kernel (multiples of 4: 100, 300) and the single-entry one (33)."""
    a = gen.rmat(12, seed=4)
    rng = np.random.default_rng(2)
    b = rng.standard_normal((a.n, 104)).astype(np.float32)
    c0 = rng.standard_normal((a.m, 104)).astype(np.float32)
    check(handle, a, b, c0, n=100, alpha=2.0, beta=0.5, bf16=bf16, ptr_dtype=torch.int64)
    b = rng.standard_normal((a.n, 33)).astype(np.float32)
    c0 = rng.standard_normal((a.m, 33)).astype(np.float32)
    check(handle, a, b, c0, alpha=-1.0, beta=0.0, bf16=bf16)
    b = rng.standard_normal((a.n, 300)).astype(np.float32)
    c0 = rng.standard_normal((a.m, 300)).astype(np.float32)
    check(handle, a, b, c0, bf16=bf16)


@pytest.mark.parametrize("bf16", [False, True])
def test_long_rows_cross_chunks(handle, bf16):
    """Rows longer than a wavefront's chunk are added with atomics from several chunks; rows of one entry and
    empty rows sit between them, so pairs of entries straddle row ends."""
    m, k = 8, 5000
    rp = np.array([0, 3000, 3000, 3001, 4500, 4500, 4500, 6000, 7000], np.int64)
    rng = np.random.default_rng(3)
    cols = np.concatenate([np.sort(rng.choice(k, int(rp[i + 1] - rp[i]), replace=False)) for i in range(m)]).astype(
        np.int32)
    a = CSR(m, k, rp, cols, rng.uniform(-1, 1, int(rp[-1])).astype(np.float32))
    b = rng.standard_normal((k, 64)).astype(np.float32)
    c0 = rng.standard_normal((m, 64)).astype(np.float32)
    check(handle, a, b, c0, bf16=bf16)


def test_empty(handle):
    a = CSR.empty(10, 20)
    b = np.ones((20, 8), np.float32)
    c0 = np.ones((10, 8), np.float32)
    check(handle, a, b, c0)
