"""GPU parity at the full sizes of the benchmark configurations, through the shipped paths.

Where the oracle cannot redo a whole result in seconds, the test compares every row's entry count with the oracle's
(an exact count of the pattern of C), and the full structure and values of sampled rows: the heaviest rows of each
product, rows whose outputs reach within 64 of the last column B holds, and random rows. Properties that do not depend on
size cover the rest (a closed form for the Laplacian's square; FP64 results for the dense product).

- One SUMMA stage (k = 3, and the nine-times heavier k = 2) of R-MAT scale 22 (edge factor 16) on an 8 x 8 grid of
  524288^2 tiles (2^19 columns: the width of the sort keys and the heavy-row bitmap, whose last columns the thin
  products of tests/test_wide_tiles_gpu.py fill), batched learned pipeline with
  the trained predictor and the autotuned knobs, one graph launch (bcl/containers/algorithms/spgemm.hpp:181-242).
- SpMM: R-MAT scale 20 times a dense 2^20 x 128 B, fp32 and bf16 B (spgemm.hpp:78-96).
- Dense GEMM 8192^3 in the three MFMA modes against the FP64 product (cblas_wrapper.hpp:16-40).
- FP64 SpGEMM of the 5-point Laplacian on a 4096 x 4096 grid: nnz(C) = 218,021,892 and every value by closed form.

Structure is bit-exact; sparse values are within one rounding of the float64 oracle
(tolerance of examples/experimental/array/sp_sp_matrix.cpp:122-155 is 1e-5 relative; this bound is far inside it).
"""
import os

import numpy as np
import pytest
import torch

import oracle_py as O
from bcl_amd import autotune, gen, predictor
from bcl_amd.csr import CSR
from bcl_amd.device import (GEMM_BF16, GEMM_FP32, GEMM_SPLIT_BF16, DeviceCSR, gemm, gemm_f64_ref, spgemm, spmm)
from bcl_amd.stage import LEARNED, StagePlan, d2h, stage_tiles
from conftest import spgemm_value_bound

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
NTH = min(16, os.cpu_count() or 1)  # host threads of the oracle (the GPU box gives a process 16 cores)


def _sub_rows(a: CSR, rows: np.ndarray, dtype=None) -> CSR:
    lens = np.diff(a.row_ptr)[rows]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.arange(a.row_ptr[r], a.row_ptr[r + 1]) for r in rows]) if rp[-1] else np.zeros(0, int)
    vals = a.vals[idx]
    if dtype is not None:
        vals = vals.astype(dtype)
    return CSR(len(rows), a.n, rp, a.col_idx[idx].astype(np.int32), vals)


def _with_vals(a: CSR, vals) -> CSR:
    return CSR(a.m, a.n, a.row_ptr, a.col_idx, vals)


def row_products(a: CSR, b: CSR) -> np.ndarray:
    cum = np.concatenate([[0], np.cumsum(np.diff(b.row_ptr)[a.col_idx], dtype=np.int64)])
    return cum[a.row_ptr[1:]] - cum[a.row_ptr[:-1]]


def sample_rows(a: CSR, b: CSR, n_heavy: int, n_edge: int, n_rand: int, seed: int) -> np.ndarray:
    """The n_heavy rows with the most products, n_edge random rows whose output reaches within 64 of the largest
    column B holds (R-MAT leaves the last columns of a tile empty: at scale 22 the largest held column of a 2^19-wide
    tile is about 2^19 - 250), and n_rand random non-empty rows."""
    rng = np.random.default_rng(seed)
    prod = row_products(a, b)
    heavy = np.argsort(prod, kind="stable")[-n_heavy:]
    blen = np.diff(b.row_ptr)
    last = np.where(blen > 0, b.col_idx[np.maximum(b.row_ptr[1:] - 1, 0)], -1)
    reach = (last >= int(last.max()) - 64).astype(np.int64)
    cum = np.concatenate([[0], np.cumsum(reach[a.col_idx])])
    edge_rows = np.nonzero(cum[a.row_ptr[1:]] - cum[a.row_ptr[:-1]] > 0)[0]
    edge = rng.choice(edge_rows, min(n_edge, edge_rows.size), replace=False) if edge_rows.size else edge_rows
    nz = np.nonzero(prod > 0)[0]
    rand = rng.choice(nz, min(n_rand, nz.size), replace=False) if nz.size else nz
    return np.unique(np.concatenate([heavy, edge, rand]).astype(np.int64)), edge_rows.size


def gpu_rows(desc, row_ptr: np.ndarray, rows: np.ndarray):
    """Columns and values of the given rows of a device result (one copy per row)."""
    cols, vals = [], []
    vdt = np.float32 if desc.dtype == 0 else np.float64
    for r in rows:
        lo, hi = int(row_ptr[r]), int(row_ptr[r + 1])
        c = np.zeros(hi - lo, np.int32)
        v = np.zeros(hi - lo, vdt)
        if hi > lo:
            d2h(c, desc.col_idx + 4 * lo)
            d2h(v, desc.vals + v.itemsize * lo)
        cols.append(c)
        vals.append(v)
    return cols, vals


def check_product_sampled(plan, p, a, b, n_heavy, seed):
    d = plan.result_desc(p)
    assert d.m == a.m and d.n == b.n
    rp = np.zeros(a.m + 1, np.int64)
    d2h(rp, d.row_ptr)
    # every row's entry count against the oracle's exact count of the pattern of C
    want_nnz, _ = O.row_nnz(a, b, nthreads=NTH)
    assert rp[0] == 0 and np.array_equal(np.diff(rp), want_nnz), f"product {p}: row counts differ"
    rows, n_edge_rows = sample_rows(a, b, n_heavy, 100, 100, seed)
    sub = _sub_rows(a, rows, np.float64)
    b64 = _with_vals(b, b.vals.astype(np.float64))
    ref = O.spgemm_rowhash(sub, b64, nthreads=NTH)  # structure, and values summed in float64 in A-row order
    absum = O.spgemm_rowhash(_with_vals(sub, np.abs(sub.vals)), _with_vals(b64, np.abs(b64.vals)), nthreads=NTH)
    cols, vals = gpu_rows(d, rp, rows)
    gcol = np.concatenate(cols)
    gval = np.concatenate(vals).astype(np.float64)
    np.testing.assert_array_equal(np.diff(ref.row_ptr), np.diff(rp)[rows])
    np.testing.assert_array_equal(gcol, ref.col_idx)
    err = np.abs(gval - ref.vals)
    bound = spgemm_value_bound(sub, b, ref.vals, absum.vals, np.float32)
    assert (err <= bound).all(), f"product {p}: {(err > bound).sum()} values out of bound"
    b_last = int(b.col_idx.max()) if b.nnz else -1
    if n_edge_rows:
        assert gcol.max() >= b_last - 64, f"product {p}: the sampled rows must reach the last columns B holds"
    return {"rows": int(rows.size), "entries": int(gcol.size), "edge_rows": int(n_edge_rows),
            "max_rel": float((err / np.maximum(np.abs(ref.vals), 1e-300)).max()) if err.size else 0.0,
            "max_col": int(gcol.max()) if gcol.size else -1, "b_max_col": b_last}


@pytest.fixture(scope="module")
def rmat22_tiles():
    a = gen.rmat(22, ef=16, seed=1234)
    tiles = gen.split_grid(a, 8, 8)
    del a
    return tiles


@pytest.mark.parametrize("k,n_heavy", [(3, 100), (2, 50)])
def test_stage_rmat22_grid8(handle, rmat22_tiles, k, n_heavy):
    A, B, pairs = stage_tiles(rmat22_tiles, rmat22_tiles, 8, k)
    assert all(b.n == 1 << 19 for _, b in pairs)
    tun, margin, slack = autotune.load_tuning()
    print(f"[s22] stage k={k}: learned plan (shipped knobs {None if tun is None else list(tun)})", flush=True)
    plan = StagePlan(handle, A, B, mode=LEARNED, params=predictor.load(), margin=margin, slack=slack, tuning=tun)
    plan.execute()
    plan.execute()  # a graph replay: the result checked is the replay's
    torch.cuda.synchronize()
    st = plan.stats()
    print(f"[s22] stats {st}", flush=True)
    edge = 0
    for p, (a, b) in enumerate(pairs):
        info = check_product_sampled(plan, p, a, b, n_heavy, seed=100 * k + p)
        edge += info["edge_rows"] > 0
        if p % 8 == 0:
            print(f"[s22] k={k} product {p}: {info}", flush=True)
    assert edge >= len(pairs) // 2, "most products must have sampled rows at the right edge of their columns"
    assert st["heavy_rows"] > 0 and st["sorted_rows"] > 0
    plan.close()
    del A, B
    torch.cuda.empty_cache()


def test_spmm_rmat20_x_dense128(handle):
    a = gen.rmat(20, ef=16, seed=1234)
    n = 128
    A = DeviceCSR.from_host(a)
    g = torch.Generator(device="cuda").manual_seed(7)
    Bf = torch.randn((a.n, n), device="cuda", dtype=torch.float32, generator=g)
    lens = np.diff(a.row_ptr)
    rng = np.random.default_rng(3)
    rows = np.unique(np.concatenate([np.argsort(lens)[-100:], rng.choice(a.m, 2048, replace=False)]))
    sub = _sub_rows(a, rows)
    for name, Bt in (("fp32", Bf), ("bf16", Bf.to(torch.bfloat16))):
        C = torch.full((a.m, n), 0.25, device="cuda", dtype=torch.float32)  # C += A B keeps what C held
        spmm(handle, A, Bt, C)
        torch.cuda.synchronize()
        got = C[torch.from_numpy(rows).cuda()].double().cpu().numpy()
        bh = Bt.view(torch.int16).cpu().numpy().view(np.uint16) if name == "bf16" else Bt.cpu().numpy()
        ref = O.spmm_f64(sub, bh, c0=np.full((rows.size, n), 0.25))
        bval = (bh.astype(np.uint32) << 16).view(np.float32) if name == "bf16" else bh
        absum = O.spmm_f64(_with_vals(sub, np.abs(sub.vals)), np.abs(bval)) + 0.25
        L = np.diff(sub.row_ptr)[:, None] + 2
        bound = (L * U32 / (1 - L * U32)) * absum
        err = np.abs(got - ref)
        assert (err <= bound).all(), f"{name}: {(err > bound).sum()} entries out of the fp32 summation bound"
        rel = err / np.maximum(absum, 1e-300)
        print(f"[spmm] {name} B: error / sum|a||b| median {np.median(rel):.3e} p99 {np.quantile(rel, 0.99):.3e} "
              f"max {rel.max():.3e}", flush=True)
        if name == "bf16":  # what rounding B to bf16 costs against the fp32-B product in FP64
            ref32 = O.spmm_f64(sub, Bf.cpu().numpy(), c0=np.full((rows.size, n), 0.25))
            d = np.abs(got - ref32) / np.maximum(absum, 1e-300)
            print(f"[spmm] bf16 B vs fp32-B FP64: median {np.median(d):.3e} p99 {np.quantile(d, 0.99):.3e}", flush=True)
            assert np.quantile(d, 0.99) < 2.0 ** -8
        del C


def test_gemm_8192_three_modes(handle):
    N = 8192
    g = torch.Generator(device="cuda").manual_seed(11)
    A = torch.randn((N, N), device="cuda", generator=g)
    B = torch.randn((N, N), device="cuda", generator=g)
    C64 = gemm_f64_ref(handle, A, B)
    rows = np.arange(0, N, N // 16)
    ref_rows = O.gemm_f64(A[torch.from_numpy(rows).cuda()].cpu().numpy(), B.cpu().numpy())
    np.testing.assert_allclose(C64[torch.from_numpy(rows).cuda()].cpu().numpy(), ref_rows, rtol=1e-12, atol=1e-9)
    S = gemm_f64_ref(handle, A.abs(), B.abs())  # sum_k |a_ik| |b_kj|, exact in float64
    Smax = float(S.max())
    K = N
    bounds = {GEMM_FP32: (K + 1) * U32 * 1.01,
              GEMM_BF16: 2.0 ** -8 + 2.0 ** -18 + K * U32 * 1.01,
              GEMM_SPLIT_BF16: 3.01 * 2.0 ** -18 + 3 * K * U32 * 1.02}
    res = {}
    for mode, name in ((GEMM_FP32, "fp32"), (GEMM_BF16, "bf16"), (GEMM_SPLIT_BF16, "split_bf16_3")):
        C = torch.zeros((N, N), device="cuda")
        gemm(handle, A, B, C, mode=mode, beta=0.0)
        torch.cuda.synchronize()
        err = (C.double() - C64).abs()
        assert bool((err <= bounds[mode] * S).all()), f"{name}: entries beyond the rounding bound"
        rel = (err / S).flatten()
        q = torch.quantile(rel[torch.randint(0, rel.numel(), (1 << 20,), device="cuda",
                                             generator=torch.Generator(device="cuda").manual_seed(5))],
                           torch.tensor([0.5, 0.99], device="cuda", dtype=torch.float64))
        res[name] = (float(q[0]), float(q[1]), float(err.max()) / Smax)
        print(f"[gemm] {name}: error / sum|a||b| median {res[name][0]:.3e} p99 {res[name][1]:.3e}; "
              f"max error / max sum|a||b| {res[name][2]:.3e}", flush=True)
        del C, err, rel
    # split-bf16 recovers fp32 accuracy: hi + lo represents each input to 2^-18 relative (the lo part's own rounding),
    # so its typical error is within a small factor of the fp32 product's (about 2x at K = 8192), its largest error
    # that of fp32, and both far below bf16's
    assert res["split_bf16_3"][1] <= 4.0 * res["fp32"][1] + 1e-9
    assert res["split_bf16_3"][2] <= 2.0 * res["fp32"][2] + 1e-9
    assert res["split_bf16_3"][1] < 0.05 * res["bf16"][1]


def test_laplacian4096_fp64_closed_form(handle):
    gsz = 4096
    a = gen.laplacian2d(gsz)
    A = DeviceCSR.from_host(a)
    C = spgemm(handle, A, A)
    torch.cuda.synchronize()
    assert C.nnz == 218_021_892
    rp = C.row_ptr
    cnt = rp[1:] - rp[:-1]
    m = a.m
    i = torch.arange(m, device="cuda", dtype=torch.int64)
    xi, yi = i // gsz, i % gsz
    # entries of row i: grid points within Manhattan distance 2
    want = torch.zeros(m, dtype=torch.int64, device="cuda")
    deg = torch.zeros(m, dtype=torch.int64, device="cuda")
    for dx in range(-2, 3):
        for dy in range(-2, 3):
            if abs(dx) + abs(dy) > 2:
                continue
            ok = (xi + dx >= 0) & (xi + dx < gsz) & (yi + dy >= 0) & (yi + dy < gsz)
            want += ok.long()
            if abs(dx) + abs(dy) == 1:
                deg += ok.long()
    assert torch.equal(cnt, want), "row entry counts differ from the stencil's"
    row = torch.repeat_interleave(i, cnt)
    col = C.col_idx.long()
    dxe, dye = col // gsz - row // gsz, col % gsz - row % gsz
    man = dxe.abs() + dye.abs()
    expect = torch.where(man == 0, 16.0 + deg[row].double(),
                         torch.where(man == 1, -8.0, torch.where(dxe.abs() == 1, 2.0, 1.0)))
    assert bool((man <= 2).all())
    assert torch.equal(C.vals, expect), "values differ from the closed form of A^2"
    del row, col, dxe, dye, man, expect
    # columns strictly increasing within each row
    dc = C.col_idx[1:].long() - C.col_idx[:-1].long()
    starts = torch.zeros(C.nnz, dtype=torch.bool, device="cuda")
    starts[rp[1:-1][cnt[1:] > 0]] = True
    assert bool((dc[~starts[1:]] > 0).all())
