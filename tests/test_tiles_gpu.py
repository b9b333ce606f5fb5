"""GPU parity of the tile ingest and of DMatrix's elementwise tile arithmetic.

- bclx_csr_split_grid against the reference's own CSRMatrix::get_slice_impl_ (oracle/_ref, when built) and the host
  tile cutter, bit for bit: ragged edge tiles, empty rows and empty tiles, int32 and int64 row_ptr, fp32 and fp64
  values, a grid of one tile, an empty matrix; at full size, the R-MAT scale-22 matrix (65 M entries) cut into the
  8 x 8 grid of the batched stage, against the host cut. Columns out of order are refused by the kernel's own checks.
- bclx_dense_binary_op (DMatrix +, -, *, / and +=, bcl/containers/DMatrix.hpp:319-386) against numpy, bit for bit
  (IEEE-rounded elementwise operations), with padded leading dimensions, in place, and on denormals and infinities.
"""
import numpy as np
import pytest
import torch

import oracle_py as O
from bcl_amd import gen
from bcl_amd._lib import BclxError
from bcl_amd.csr import CSR, get_slice
from bcl_amd.device import (OP_ADD, OP_DIV, OP_MUL, OP_SUB, DeviceCSR, dense_binary_op, split_grid)

pytestmark = pytest.mark.gpu


def host_tiles(a: CSR, tm: int, tn: int):
    gm, gn = max(1, -(-a.m // tm)), max(1, -(-a.n // tn))
    return {(i, j): get_slice(a, i * tm, (i + 1) * tm, j * tn, (j + 1) * tn) for i in range(gm) for j in range(gn)}


def assert_tile(got: CSR, want: CSR, what: str):
    assert (got.m, got.n, got.nnz) == (want.m, want.n, want.nnz), what
    np.testing.assert_array_equal(got.row_ptr, want.row_ptr, what)
    np.testing.assert_array_equal(got.col_idx, want.col_idx, what)
    np.testing.assert_array_equal(got.vals.view(np.uint8), np.ascontiguousarray(want.vals).view(np.uint8), what)


def with_empty_rows(a: CSR, every: int) -> CSR:
    keep = np.ones(a.m, bool)
    keep[::every] = False
    rows = np.repeat(np.arange(a.m), a.row_nnz())
    mask = keep[rows]
    rp = np.zeros(a.m + 1, np.int64)
    rp[1:] = np.cumsum(np.bincount(rows[mask], minlength=a.m))
    return CSR(a.m, a.n, rp, a.col_idx[mask], a.vals[mask])


@pytest.mark.parametrize("tm,tn,ptr64,f64", [(300, 250, False, False), (1000, 97, True, False), (64, 4096, False, True),
                                            (5000, 5000, True, True), (1, 1024, False, False)])
def test_split_grid_matches_slices(handle, tm, tn, ptr64, f64):
    a = with_empty_rows(gen.rmat(12, seed=31), 7)  # 4096 x 4096, then its first 3000 columns
    keep = a.col_idx < 3000
    rows = np.repeat(np.arange(a.m), a.row_nnz())
    rp = np.zeros(a.m + 1, np.int64)
    rp[1:] = np.cumsum(np.bincount(rows[keep], minlength=a.m))
    a = CSR(a.m, 3000, rp, a.col_idx[keep], a.vals[keep].astype(np.float64 if f64 else np.float32))
    pd = torch.int64 if ptr64 else torch.int32
    print(f"[split] {a.m}x{a.n} nnz={a.nnz} tiles {tm}x{tn} ptr64={ptr64} f64={f64}", flush=True)
    A = DeviceCSR.from_host(a, ptr_dtype=pd)
    got = split_grid(handle, A, tm, tn, ptr_dtype=pd)
    torch.cuda.synchronize()
    want = host_tiles(a, tm, tn)
    assert set(got) == set(want)
    for key, w in want.items():
        assert_tile(got[key].to_host(), w, f"tile {key}")
        if O.have_ref() and not f64 and key[0] % 3 == 0:
            i, j = key
            assert_tile(got[key].to_host(), O.ref_slice(a, i * tm, (i + 1) * tm, j * tn, (j + 1) * tn), f"ref {key}")


def test_split_grid_empty_and_single(handle):
    e = CSR.empty(10, 20)
    t = split_grid(handle, DeviceCSR.from_host(e), 4, 8)
    assert len(t) == 3 * 3 and all(x.nnz == 0 for x in t.values())
    assert all(int(x.row_ptr.abs().sum()) == 0 and x.row_ptr.numel() == x.m + 1 for x in t.values())
    z = split_grid(handle, DeviceCSR.from_host(CSR.empty(0, 5)), 4, 4)  # no rows: one row of two tiles
    assert sorted(z) == [(0, 0), (0, 1)] and [z[k].n for k in sorted(z)] == [4, 1]
    assert all(x.m == 0 and x.row_ptr.numel() == 1 and int(x.row_ptr[0]) == 0 for x in z.values())
    a = gen.uniform(100, 80, 5, seed=2)
    one = split_grid(handle, DeviceCSR.from_host(a), 100, 80)
    assert_tile(one[(0, 0)].to_host(), a, "one tile")


def test_split_grid_refuses_unsorted_columns(handle):
    a = gen.uniform(64, 64, 4, seed=4)
    col = a.col_idx.copy()
    col[8:12] = col[8:12][::-1]  # row 2 out of order
    A = DeviceCSR.from_host(CSR(a.m, a.n, a.row_ptr, col, a.vals))
    with pytest.raises(BclxError) as e:
        split_grid(handle, A, 16, 16)
    assert e.value.status == 1 and "sorted" in str(e.value)


def test_split_grid_rmat22_grid8(handle):
    a = gen.rmat(22, ef=16, seed=1234)
    want = gen.split_grid(a, 8, 8)
    A = DeviceCSR.from_host(a)
    tm = -(-a.m // 8)
    print(f"[split] rmat22 nnz={a.nnz}", flush=True)
    got = split_grid(handle, A, tm, tm)
    torch.cuda.synchronize()
    del A
    for key, w in want.items():
        g = got[key]
        assert (g.m, g.n, g.nnz) == (w.m, w.n, w.nnz), key
        assert torch.equal(g.row_ptr.cpu().to(torch.int64), torch.from_numpy(w.row_ptr.astype(np.int64))), key
        assert torch.equal(g.col_idx.cpu(), torch.from_numpy(w.col_idx)), key
        assert torch.equal(g.vals.cpu(), torch.from_numpy(w.vals)), key


OPS = {OP_ADD: np.add, OP_SUB: np.subtract, OP_MUL: np.multiply, OP_DIV: np.divide}


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("op", sorted(OPS))
@pytest.mark.parametrize("shape,ld", [((257, 129), None), ((64, 100), 128), ((1000, 1000), None), ((3, 5), 7)])
def test_dense_binary_op(handle, dtype, op, shape, ld):
    m, n = shape
    g = torch.Generator().manual_seed(op * 7 + m)
    ldx = ld or n
    A = torch.randn((m, ldx), generator=g, dtype=dtype)
    B = torch.randn((m, ldx), generator=g, dtype=dtype)
    tiny = torch.finfo(dtype).tiny
    A[0, :4] = torch.tensor([tiny / 8, -tiny / 4, float("inf"), 0.0], dtype=dtype)  # denormals, inf, zero
    B[0, :4] = torch.tensor([tiny / 16, 3.0, 2.0, 5.0], dtype=dtype)
    want = OPS[op](A[:, :n].numpy(), B[:, :n].numpy())
    dA, dB = A.cuda(), B.cuda()
    dC = torch.full((m, ldx), 7.0, dtype=dtype, device="cuda")
    dense_binary_op(handle, op, dA[:, :n], dB[:, :n], dC[:, :n])
    got = dC.cpu()
    np.testing.assert_array_equal(got[:, :n].numpy().view(np.uint8), want.view(np.uint8))
    assert bool((got[:, n:] == 7.0).all()), "padding columns must be left alone"
    dense_binary_op(handle, op, dA[:, :n], dB[:, :n])  # in place: A op= B
    np.testing.assert_array_equal(dA.cpu()[:, :n].numpy().view(np.uint8), want.view(np.uint8))
