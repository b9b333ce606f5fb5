"""GPU parity of the register-sorted rows: rows of at most 32 products, half a wavefront each.

Their sort keys pack the column in 27 bits and the product index in 5, so they take any tile narrower than
2^27 - 1 columns (the 5-point Laplacian's square on a 4096^2 grid is 16.7 M columns wide, 25 products per row).
Cases: rows of 1 to 32 products with repeated columns (runs summed in A-row order), rows of more than 32 A entries
whose B rows are mostly empty (the entries are staged 32 at a time), empty rows, columns at the widest limit, and
one column past it (those rows hash instead). Through the two-phase call (symbolic then numeric), the batched exact
stage and the batched learned stage, in fp32 and fp64 with int32 and int64 row_ptr. Structure must equal the
oracle's (bcl/containers/algorithms/spgemm.hpp:40-76) and values be within one rounding of the float64 oracle.
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


def small_product(n: int, dtype, seed: int):
    """A: 2000 x 3000, B: 3000 x n. B rows 0..1999 hold 0 to 4 columns drawn from a pool of 24 (so products of
    one A row repeat columns), some reaching n - 1; B rows 2000..2999 are empty. A rows: 1 to 8 entries on the
    non-empty B rows, or 40 to 60 entries of which at most 6 are non-empty B rows, or none."""
    rng = np.random.default_rng(seed)
    pool = np.unique(np.concatenate([rng.choice(n, 20, replace=False), [n - 1, n - 2, 0, 1]]))
    brows = [np.sort(rng.choice(pool, int(rng.integers(0, 5)), replace=False)) for _ in range(2000)]
    brows += [np.zeros(0, np.int64)] * 1000
    arows = []
    for r in range(2000):
        kind = r % 5
        if kind == 4:
            arows.append(np.zeros(0, np.int64))
        elif kind == 3:
            live = rng.choice(2000, int(rng.integers(1, 7)), replace=False)
            dead = rng.choice(np.arange(2000, 3000), int(rng.integers(40, 61)) - live.size, replace=False)
            arows.append(np.sort(np.concatenate([live, dead])))
        else:
            arows.append(np.sort(rng.choice(2000, int(rng.integers(1, 9)), replace=False)))

    def mk(rows, ncols):
        rp = np.concatenate([[0], np.cumsum([len(x) for x in rows])]).astype(np.int64)
        col = np.concatenate(rows).astype(np.int32)
        return CSR(len(rows), ncols, rp, col, rng.uniform(-1, 1, col.size).astype(dtype))

    return mk(arows, 3000), mk(brows, n)


@pytest.mark.parametrize("n,f64,ptr64", [(1 << 20, False, False), ((1 << 27) - 2, False, True),
                                         ((1 << 27) - 1, True, False), (16_777_216, True, True)])
def test_small_rows(handle, n, f64, ptr64):
    a, b = small_product(n, np.float64 if f64 else np.float32, seed=n % 1009)
    blen = np.diff(b.row_ptr.astype(np.int64))
    prod = np.array([blen[a.col_idx[a.row_ptr[i]:a.row_ptr[i + 1]]].sum() for i in range(a.m)])
    assert prod.max() <= 32 and (prod > 0).sum() > 1000
    assert (np.diff(a.row_ptr) > 32).sum() > 300, "rows of more than 32 entries must be staged in chunks"
    pd = torch.int64 if ptr64 else torch.int32
    print(f"[small] n={n} f64={f64} ptr64={ptr64} A nnz={a.nnz} B nnz={b.nnz} rows with products "
          f"{int((prod > 0).sum())}", flush=True)
    A, B = DeviceCSR.from_host(a, ptr_dtype=pd), DeviceCSR.from_host(b, ptr_dtype=pd)
    ref = _check(spgemm(handle, A, B, ptr_dtype=pd).to_host(), a, b, f"two-phase n={n}")
    assert ref.col_idx.max() >= n - 2, "outputs must reach the last columns"
    assert (np.diff(ref.row_ptr) < prod).any(), "some rows must repeat columns"
    tun, margin, slack = autotune.load_tuning()
    for mode in (BATCHED_EXACT, LEARNED):
        print(f"[small] n={n} stage mode {mode}", flush=True)
        plan = StagePlan(handle, [A], [B], mode=mode, params=predictor.load(), margin=margin, slack=slack,
                         tuning=tun)
        plan.execute()
        torch.cuda.synchronize()
        st = plan.stats()
        print("[small] stats", st, flush=True)
        _check(plan.result_host(0), a, b, f"stage mode {mode} n={n}")
        expect = int((prod > 0).sum()) if n < (1 << 27) - 1 else 0
        assert st["sorted_rows"] == expect, "rows of at most 32 products sort in registers below 2^27 - 1 columns"
        plan.close()


def test_laplacian_rows_small_bin(handle):
    """The Laplacian's rows (25 products, 13 outputs) on a grid wider than the ESC key: bit-exact integers."""
    a = gen.laplacian2d(200)
    a = CSR(a.m, a.n, a.row_ptr, a.col_idx, a.vals.astype(np.float64))
    c = spgemm(handle, DeviceCSR.from_host(a), DeviceCSR.from_host(a)).to_host()
    ref = O.spgemm_f64(a, a)
    np.testing.assert_array_equal(c.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(c.col_idx, ref.col_idx)
    np.testing.assert_array_equal(c.vals, ref.vals)
