"""GPU parity of the batched SUMMA stage (bclx_stage_*) against the oracle, in all three modes.

Every product of the stage must have the oracle's exact structure and values within one rounding of the float64
oracle, whether it ran plain, batched-exact, or batched-learned (with a trained, an untrained or no predictor: an
untrained predictor makes most rows overflow their predicted slot and exercises the exact redo).
"""
import numpy as np
import pytest
import torch

import oracle_py as O
from bcl_amd import autotune, gen, predictor
from bcl_amd.csr import CSR
from bcl_amd.stage import BATCHED_EXACT, LEARNED, PLAIN, StagePlan, stage_tiles
from conftest import spgemm_value_bound

pytestmark = pytest.mark.gpu


def _abs(a: CSR) -> CSR:
    return CSR(a.m, a.n, a.row_ptr, a.col_idx, np.abs(a.vals))


@pytest.fixture(scope="module")
def stage14():
    """R-MAT scale 14 on a 2 x 2 grid (8192^2 tiles, hub rows need the global accumulator), stage k = 0."""
    a = gen.rmat(14, seed=31)
    tiles = gen.split_grid(a, 2, 2)
    return tiles


@pytest.fixture(scope="module")
def stage13():
    """R-MAT scale 13 on a 4 x 4 grid (2048^2 tiles, very uneven), stage k = 1."""
    a = gen.rmat(13, seed=33)
    return gen.split_grid(a, 4, 4)


def check_stage(plan: StagePlan, pairs, scale=1.0):
    for p, (a, b) in enumerate(pairs):
        c = plan.result_host(p)
        ref = O.spgemm(a, b)
        assert c.nnz == ref.nnz, f"product {p}: nnz {c.nnz} != {ref.nnz}"
        np.testing.assert_array_equal(c.row_ptr, ref.row_ptr)
        np.testing.assert_array_equal(c.col_idx, ref.col_idx)
        c64 = O.spgemm_f64(a, b)
        absum = O.spgemm_f64(_abs(a), _abs(b))
        err = np.abs(c.vals.astype(np.float64) - scale * c64.vals)
        bound = spgemm_value_bound(a, b, scale * c64.vals, scale * absum.vals, np.float32)
        assert (err <= bound).all(), f"product {p}: {(err > bound).sum()} values out of bound"


def _params(kind):
    if kind is None:
        return None
    if kind == "untrained":
        return predictor.init_params(seed=5)
    p = predictor.load()
    if p is None:
        pytest.skip("no trained predictor weights in bcl_amd/data")
    return p


@pytest.mark.parametrize("mode,pkind,graph", [
    (PLAIN, None, False),
    (BATCHED_EXACT, None, True),
    (BATCHED_EXACT, None, False),
    (LEARNED, None, True),
    (LEARNED, "untrained", True),
    (LEARNED, "untrained", False),
    (LEARNED, "trained", True),
])
def test_stage_modes(handle, stage14, mode, pkind, graph):
    print(f"[stage] mode={mode} params={pkind} graph={graph}", flush=True)
    A, B, pairs = stage_tiles(stage14, stage14, 2, 0)
    plan = StagePlan(handle, A, B, mode=mode, params=_params(pkind), use_graph=graph)
    plan.execute()
    torch.cuda.synchronize()
    st = plan.stats()
    print("[stage] stats", st, flush=True)
    check_stage(plan, pairs)
    if mode != PLAIN:
        assert st["heavy_rows"] > 0, "this stage must exercise the global accumulator"
    if pkind == "untrained":
        assert st["overflow_rows"] > 0, "an untrained predictor must exercise the overflow redo"
    # replay: same results
    plan.execute()
    torch.cuda.synchronize()
    check_stage(plan, pairs)
    plan.close()


@pytest.mark.parametrize("tuning", [(64, 4096, 1, 32, 0), (16, 1024, 2, 0, 4096), (256, 2048, 1, 64, 256),
                                    (64, 4096, 1, 32, 1024), (16, 2048, 1, 16, 4096, 2), (64, 4096, 1, 32, 4096, 4)])
def test_stage_tuning_choices(handle, stage14, tuning):
    """Every binning choice (sorting off, all sortable rows sorted, narrow groups, bins run as two or four parallel
    graph branches) gives the oracle's result."""
    print(f"[stage] tuning={tuning}", flush=True)
    A, B, pairs = stage_tiles(stage14, stage14, 2, 1)
    for mode in (BATCHED_EXACT, LEARNED):
        plan = StagePlan(handle, A, B, mode=mode, params=predictor.init_params(seed=3), tuning=np.array(tuning))
        plan.execute()
        torch.cuda.synchronize()
        check_stage(plan, pairs)
        plan.close()


def test_stage_sliced_heavy_rows(handle):
    """One product of R-MAT scale 15 (32768 columns): heavy rows that span several LDS slices, in the predicted mode
    (slot or bump allocation) and the one-pass redo of overflowed rows (untrained network)."""
    a = gen.rmat(15, seed=7)
    tiles = gen.split_grid(a, 1, 1)
    A, B, pairs = stage_tiles(tiles, tiles, 1, 0)
    for prm in (predictor.load(), predictor.init_params(seed=4)):
        if prm is None:
            continue
        plan = StagePlan(handle, A, B, mode=LEARNED, params=prm)
        plan.execute()
        torch.cuda.synchronize()
        check_stage(plan, pairs)
        plan.close()


def test_stage_uneven_grid_learned(handle, stage13):
    A, B, pairs = stage_tiles(stage13, stage13, 4, 1)
    for mode in (BATCHED_EXACT, LEARNED):
        plan = StagePlan(handle, A, B, mode=mode, params=predictor.init_params(seed=1))
        plan.execute()
        torch.cuda.synchronize()
        check_stage(plan, pairs)
        plan.close()


def test_stage_replay_with_new_values(handle, stage13):
    """A plan is bound to the sparsity structure; replaying the graph after the values change recomputes them."""
    A, B, pairs = stage_tiles(stage13, stage13, 4, 2)
    plan = StagePlan(handle, A, B, mode=LEARNED, params=predictor.init_params(seed=2))
    plan.execute()
    torch.cuda.synchronize()
    check_stage(plan, pairs)
    seen = set()
    for t in A + B:
        if id(t) not in seen:
            t.vals.mul_(2.0)
            seen.add(id(t))
    plan.execute()
    torch.cuda.synchronize()
    check_stage(plan, pairs, scale=4.0)
    plan.close()


def _median_ms(h, plan, reps=7):
    plan.execute()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(h.stream)
        plan.execute()
        e.record(h.stream)
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e))
    return float(np.median(ts))


@pytest.mark.parametrize("sorting", ["off", "shipped"])
@pytest.mark.parametrize("kind", ["banded", "rmat"])
def test_predictor_gain(handle, kind, sorting):
    """What the trained network gains over the method without it, on the same stage.

    Without the network every row's capacity is the exact bound min(products, n); the exact pipeline runs the
    symbolic pass instead. All three must give the oracle's result; the times and slot capacities are reported.
    With sorting off every row hashes and the network sizes all of them: it must size the slot region well below the
    bound (its point: smaller tables and output arena) with few overflows. With the shipped tuning (sorting on, the
    autotuned knobs and margin) sorted rows are exact in one pass and the network sizes only the rows beyond the sort
    bins: its capacities must not exceed the bound.
    """
    params = predictor.load()
    if params is None:
        pytest.skip("no trained predictor weights in bcl_amd/data")
    if sorting == "shipped":
        tun, margin, slack = autotune.load_tuning()
        if tun is None:
            pytest.skip("no autotuned settings in bcl_amd/data")
        opts = dict(tuning=tun, margin=margin, slack=slack)
    else:
        opts = dict(tuning=np.array([64, 4096, 1, 32, 0]))
    if kind == "banded":
        a = gen.banded(1 << 15, 1 << 15, bw=64, max_per_row=24, seed=11)
        grid, k = 2, 0
    else:
        a = gen.rmat(15, seed=12)
        grid, k = 4, 1
    tiles = gen.split_grid(a, grid, grid)
    A, B, pairs = stage_tiles(tiles, tiles, grid, k)
    res = {}
    for name, mode, prm in (("trained", LEARNED, params), ("bound", LEARNED, None), ("exact", BATCHED_EXACT, None)):
        print(f"[gain] {kind} sorting {sorting}: {name}", flush=True)
        plan = StagePlan(handle, A, B, mode=mode, params=prm, **opts)
        ms = _median_ms(handle, plan)
        st = plan.stats()
        check_stage(plan, pairs)
        res[name] = (ms, st)
        plan.close()
    t, b, x = res["trained"], res["bound"], res["exact"]
    nnz = t[1]["nnz"]
    assert b[1]["nnz"] == nnz and x[1]["nnz"] == nnz
    print(f"[gain] {kind} sorting {sorting}: nnz={nnz} products={t[1]['products']} | trained {t[0]:.3f} ms, slots "
          f"{t[1]['capacity']}, overflow {t[1]['overflow_rows']}/{t[1]['predicted_rows']} | bound {b[0]:.3f} ms, "
          f"slots {b[1]['capacity']} | exact {x[0]:.3f} ms", flush=True)
    if sorting == "shipped":
        assert t[1]["capacity"] <= b[1]["capacity"]
        return
    assert t[1]["predicted_rows"] > 0
    assert t[1]["overflow_rows"] <= 0.02 * t[1]["predicted_rows"] + 8
    assert t[1]["capacity"] < b[1]["capacity"]
    if kind == "banded":  # products / nnz is about 1.5 here; the trained capacities carry a 15% margin
        assert t[1]["capacity"] < 0.9 * b[1]["capacity"]
