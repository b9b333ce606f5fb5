"""GPU parity at the width limits of the accumulators and for the wider index and value types.

Tile widths around the limits the kernels are built for, on thin products (a few hundred rows of A with hub rows of
tens of thousands of products, B rows spread over the whole width and reaching its last 64 columns):
- n = 2^19 exactly: the sort keys' column field (19 bits) and the heavy-row LDS bitmap (16384 words) are full;
- n = 2^19 + 1: sorting is off (columns no longer fit the key) and heavy rows take the wide-tile kernel with its
  bitmap in LDS;
- n = 1.5 M: too wide for an LDS bitmap, so the wide-tile kernel keeps per-workgroup bitmaps in global memory.
Each runs through the two-phase call, the batched exact stage and the batched learned stage (shipped knobs).

Then R-MAT products with int64 row_ptr and float64 values, where the heavy-row kernel uses its 7168-entry slices
(several per row) and the sort kernels stage float64 products, through the two-phase call and one stage plan.
Structure must equal the oracle's (bcl/containers/algorithms/spgemm.hpp:40-76) and values be within one rounding
of the float64 oracle.
"""
import numpy as np
import pytest
import torch

import oracle_py as O
from bcl_amd import autotune, gen, predictor
from bcl_amd.csr import CSR
from bcl_amd.device import DeviceCSR, spgemm
from bcl_amd.stage import BATCHED_EXACT, LEARNED, StagePlan
from conftest import spgemm_value_bound

pytestmark = pytest.mark.gpu


def _abs(a: CSR) -> CSR:
    return CSR(a.m, a.n, a.row_ptr, a.col_idx, np.abs(a.vals))


def _check(c: CSR, a: CSR, b: CSR, what: str):
    ref = O.spgemm(a, b)
    assert c.nnz == ref.nnz, f"{what}: nnz {c.nnz} != {ref.nnz}"
    np.testing.assert_array_equal(c.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(c.col_idx, ref.col_idx)
    c64 = O.spgemm_f64(a, b)
    absum = O.spgemm_f64(_abs(a), _abs(b))
    err = np.abs(c.vals.astype(np.float64) - c64.vals)
    bound = spgemm_value_bound(a, b, c64.vals, absum.vals, a.vals.dtype)
    assert (err <= bound).all(), f"{what}: {(err > bound).sum()} values out of bound"
    return ref


def _rows_csr(rows, ncols, rng, dtype=np.float32):
    rp = np.concatenate([[0], np.cumsum([len(x) for x in rows])]).astype(np.int64)
    col = np.concatenate(rows).astype(np.int32)
    return CSR(len(rows), ncols, rp, col, rng.uniform(-1, 1, col.size).astype(dtype))


def thin_product(n: int, seed: int = 0):
    """A: 300 x 4096 (10 hub rows of 300 entries, 90 rows of 30, the rest 1 to 5); B: 4096 x n, rows of 20 to 200
    columns spread over [0, n), every 16th row also holding three of the last 64 columns."""
    rng = np.random.default_rng(seed)
    k = 4096
    brows = []
    for r in range(k):
        c = rng.choice(n, int(rng.integers(20, 200)), replace=False)
        if r % 16 == 0:
            c = np.concatenate([c, n - 1 - rng.choice(64, 3, replace=False)])
        brows.append(np.unique(c))
    arows = []
    for r in range(300):
        ln = 300 if r < 10 else 30 if r < 100 else int(rng.integers(1, 6))
        arows.append(np.sort(rng.choice(k, ln, replace=False)))
    return _rows_csr(arows, k, rng), _rows_csr(brows, n, rng)


@pytest.mark.parametrize("n", [1 << 19, (1 << 19) + 1, 1_500_000])
def test_thin_product_at_width_limits(handle, n):
    a, b = thin_product(n, seed=n % 97)
    print(f"[wide] n={n} A {a.m}x{a.n} nnz={a.nnz} B nnz={b.nnz}", flush=True)
    A, B = DeviceCSR.from_host(a), DeviceCSR.from_host(b)
    ref = _check(spgemm(handle, A, B).to_host(), a, b, f"two-phase n={n}")
    assert ref.col_idx.max() >= n - 64, "some output must reach the last 64 columns"
    tun, margin, slack = autotune.load_tuning()
    for mode in (BATCHED_EXACT, LEARNED):
        print(f"[wide] n={n} stage mode {mode}", flush=True)
        plan = StagePlan(handle, [A], [B], mode=mode, params=predictor.load(), margin=margin, slack=slack,
                         tuning=tun)
        plan.execute()
        torch.cuda.synchronize()
        st = plan.stats()
        print("[wide] stats", st, flush=True)
        _check(plan.result_host(0), a, b, f"stage mode {mode} n={n}")
        assert st["heavy_rows"] > 0, "the hub rows must take the heavy-row kernel"
        # rows of at most 32 products sort in registers at any of these widths; longer ones only when the columns
        # fit the ESC key (n <= 2^19)
        blen = np.diff(b.row_ptr.astype(np.int64))
        prod = np.array([blen[a.col_idx[a.row_ptr[i]:a.row_ptr[i + 1]]].sum() for i in range(a.m)])
        small = int(((prod > 0) & (prod <= 32)).sum())
        if n <= 1 << 19:
            assert st["sorted_rows"] > small, "rows of hundreds to thousands of products must be sorted"
        else:
            assert st["sorted_rows"] == small, "columns beyond 2^19 do not fit the ESC sort key"
        plan.close()


@pytest.mark.parametrize("scale,f64,ptr64", [(14, True, False), (15, False, True), (15, True, True)])
def test_wide_types_heavy_and_sorted_rows(handle, scale, f64, ptr64):
    a = gen.rmat(scale, seed=40 + scale)
    if f64:
        a = CSR(a.m, a.n, a.row_ptr, a.col_idx, a.vals.astype(np.float64))
    pd = torch.int64 if ptr64 else torch.int32
    print(f"[types] rmat{scale} f64={f64} ptr64={ptr64}", flush=True)
    A = DeviceCSR.from_host(a, ptr_dtype=pd)
    c = spgemm(handle, A, A, ptr_dtype=pd).to_host()
    ref = _check(c, a, a, "two-phase")
    row_nnz = np.diff(ref.row_ptr)
    assert row_nnz.max() > 7168, "a heavy row must span several 7168-entry slices"
    plan = StagePlan(handle, [A], [A], mode=LEARNED, params=predictor.load())
    plan.execute()
    torch.cuda.synchronize()
    st = plan.stats()
    print("[types] stats", st, flush=True)
    _check(plan.result_host(0), a, a, "learned stage")
    assert st["heavy_rows"] > 0 and st["sorted_rows"] > 0
    plan.close()
