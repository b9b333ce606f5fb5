"""Benchmark of the local tile multiply under BCL::gemm on one MI355X.

Default workload: one SUMMA stage of C = A * A for an R-MAT scale-22 matrix (edge factor 16, (a,b,c,d) =
(0.57, 0.19, 0.19, 0.05), duplicates merged, fp32 values) cut into an 8 x 8 grid of 524288^2 tiles: the 64
uneven products A_ik * A_kj of stage k. A step is one pass of the whole stage. The measured method is the batched
+ learned pipeline (one HIP-graph launch); the plain two-phase per-product path is timed beside it on the same
tiles, and the reference CPU tile multiply (BCL::gemm(CSRMatrix...), compiled from the reference sources when
available) is timed on a bounded row sample.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload stage|config1] [--scale 22] [--stage-k 3]

Prints one JSON line. value = useful GFLOP/s = 2 * products / time, inputs resident in HBM.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpGEMM useful GFLOP/s + HBM GB/s %peak, learned+batched vs plain GPU, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def profiled_traffic(name):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy) for this
    workload, from the committed summary profiles/<name>_hbm_traffic.md; None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{name}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def csr_bytes(m, nnz, w_ptr=4, w_val=4):
    return (m + 1) * w_ptr + nnz * (4 + w_val)


def time_steps(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    per = []
    for _ in range(steps):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record(stream)
        fn()
        e.record(stream)
        per.append((s, e))
    torch.cuda.synchronize()
    ms = [s.elapsed_time(e) for s, e in per]
    del start, end
    return float(np.sum(ms)), ms


def cpu_baseline_stage(pairs, target_products=25_000_000):
    """Reference CPU tile multiply on a bounded sample: leading rows of the heaviest product of the stage, sized to
    about `target_products` products (tens of seconds of serial CPU work)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    from bcl_amd.csr import CSR
    costs = []
    for (a, b) in pairs:
        blen = np.diff(b.row_ptr)
        costs.append(int(blen[a.col_idx].sum()) if a.nnz else 0)
    p = int(np.argmax(costs))
    a, b = pairs[p]
    blen = np.diff(b.row_ptr)
    per_entry = blen[a.col_idx].astype(np.int64)
    cum = np.concatenate([[0], np.cumsum(per_entry)])
    row_prod = cum[a.row_ptr[1:]] - cum[a.row_ptr[:-1]]
    crow = np.cumsum(row_prod)
    r = int(np.searchsorted(crow, target_products)) + 1
    r = max(1, min(r, a.m))
    sub = CSR(r, a.n, a.row_ptr[:r + 1].copy(), a.col_idx[:a.row_ptr[r]].copy(), a.vals[:a.row_ptr[r]].copy())
    prods = int(row_prod[:r].sum())
    kind = "reference" if O.have_ref() else "port"
    t = time.perf_counter()
    if kind == "reference":
        c = O.ref_spgemm(sub, b)
    else:
        c = O.spgemm(sub, b, dead_row_alloc=True)
    dt = time.perf_counter() - t
    out = {"value": round(2 * prods / dt / 1e9, 4), "unit": "GFLOP/s", "cores": 1, "kind": kind,
           "sample": f"rows 0..{r - 1} of stage product {p} (the heaviest): {prods} products, nnz(C)={c.nnz}, "
                     f"{dt:.1f} s serial BCL::gemm(CSRMatrix) incl. its per-row vector<bool>(ncols)"}
    # the same sample on every host core with the row-parallel hash restatement (SparseHashAccumulator algorithm)
    nth = min(os.cpu_count() or 1, 16)
    t = time.perf_counter()
    O.spgemm_rowhash(sub, b, nthreads=nth)
    dt2 = time.perf_counter() - t
    par = {"value": round(2 * prods / dt2 / 1e9, 4), "unit": "GFLOP/s", "cores": nth, "kind": "port",
           "sample": "same rows, row-parallel hash SpGEMM (SimpleHash per row, OpenMP)"}
    return out, par


def bench_stage(args):
    import torch
    from bcl_amd import gen, predictor
    from bcl_amd.device import Handle
    from bcl_amd.stage import LEARNED, PLAIN, BATCHED_EXACT, StagePlan, stage_tiles

    t0 = time.time()
    a = gen.rmat(args.scale, ef=16, seed=args.seed)
    log(f"R-MAT scale {args.scale}: n={a.m} nnz={a.nnz} ({time.time() - t0:.1f}s)")
    tiles = gen.split_grid(a, args.grid, args.grid)
    log(f"cut into {args.grid}x{args.grid} tiles ({time.time() - t0:.1f}s)")
    del a
    h = Handle(0)
    A, B, pairs = stage_tiles(tiles, tiles, args.grid, args.stage_k)
    stream = h.stream
    params = predictor.load()
    if params is None:
        log("WARNING: no trained predictor weights; learned mode falls back to exact upper bounds")
    tune = None
    tune_path = os.path.join(ROOT, "bcl_amd", "data", "autotune.json")
    if os.path.exists(tune_path):
        with open(tune_path) as f:
            tj = json.load(f)
        tune = np.array(tj["tuning"], np.int32)
        margin, slack = tj.get("margin", 0.15), tj.get("slack", 4)
    else:
        margin, slack = 0.15, 4

    t = time.time()
    learned = StagePlan(h, A, B, mode=LEARNED, params=params, margin=margin, slack=slack, tuning=tune)
    log(f"learned plan ({time.time() - t:.1f}s)")
    ms_l, per_l = time_steps(learned.execute, args.steps, args.warmup, stream)
    st = learned.stats()
    log("learned stats", st, "per-step ms", [round(x, 2) for x in per_l])

    pr = learned.predictions()
    pred_stats = None
    if pr["pred"].size:
        ok = pr["nnz"] > 0
        rel = np.abs(pr["pred"][ok].astype(np.float64) - pr["nnz"][ok]) / pr["nnz"][ok]
        pred_stats = {"rows": int(pr["pred"].size), "rel_err_median": round(float(np.median(rel)), 5),
                      "rel_err_p99": round(float(np.quantile(rel, 0.99)), 5),
                      "rel_err_max": round(float(rel.max()), 5),
                      "capacity_below_exact": int((pr["cap"] < pr["nnz"]).sum()),
                      "capacity_over_exact": round(float(pr["cap"].sum() / max(1, pr["nnz"].sum())), 5)}
    learned.close()

    res = {}
    if not args.skip_plain:
        t = time.time()
        plain = StagePlan(h, A, B, mode=PLAIN, tuning=tune)
        log(f"plain plan ({time.time() - t:.1f}s)")
        ms_p, per_p = time_steps(plain.execute, args.steps, args.warmup, stream)
        st_p = plain.stats()
        assert st_p["nnz"] == st["nnz"] and st_p["products"] == st["products"], (st_p, st)
        log("plain per-step ms", [round(x, 2) for x in per_p])
        plain.close()
        res["plain"] = ms_p / args.steps
    if not args.skip_exact:
        ex = StagePlan(h, A, B, mode=BATCHED_EXACT, tuning=tune)
        ms_e, per_e = time_steps(ex.execute, args.steps, args.warmup, stream)
        assert ex.stats()["nnz"] == st["nnz"]
        res["batched_exact"] = ms_e / args.steps
        ex.close()
        # the same learned pipeline without the network: every row keeps its exact upper bound min(products, n)
        bd = StagePlan(h, A, B, mode=LEARNED, params=None, margin=margin, slack=slack, tuning=tune)
        ms_b, per_b = time_steps(bd.execute, args.steps, args.warmup, stream)
        st_b = bd.stats()
        assert st_b["nnz"] == st["nnz"]
        res["bound"] = (ms_b / args.steps, st_b)
        bd.close()
        log("exact / bound per-step ms", round(res["batched_exact"], 3), round(res["bound"][0], 3))

    products = st["products"]
    flops = 2.0 * products
    ms_step = ms_l / args.steps
    gflops = flops / (ms_step * 1e-3) / 1e9
    # compulsory bytes: the 8 distinct A tiles and 8 distinct B tiles of the stage read once, every C written once
    distinct_a = {id(x): x for x in A}
    distinct_b = {id(x): x for x in B}
    in_bytes = sum(csr_bytes(x.m, x.nnz) for x in distinct_a.values()) + \
        sum(csr_bytes(x.m, x.nnz) for x in distinct_b.values())
    out_bytes = (st["rows"] + len(pairs)) * 8 + st["nnz"] * 8  # int64 row_ptr per product + (col, val)
    alg_bytes = in_bytes + out_bytes
    gbs = alg_bytes / (ms_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 values, fp64 accumulate, int32 indices",
        "data": "synthetic",
        "config": {"workload": f"rmat-s{args.scale}-ef16-grid{args.grid}-stage{args.stage_k}", "scale": args.scale,
                   "edge_factor": 16, "grid": args.grid, "stage_k": args.stage_k, "products": products,
                   "nnz_c": st["nnz"], "tiles": len(pairs), "seed": args.seed},
        "hbm_gbs": round(gbs, 2),
        "hbm_pct_peak": round(100 * gbs / HBM_PEAK_GBS, 3),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5),
                     "traffic": profiled_traffic(f"stage_rmat_s{args.scale}_grid{args.grid}_k{args.stage_k}"),
                     "kernel": "the stage's single graph launch (all pipeline kernels)",
                     "algorithmic_bytes": int(alg_bytes)},
        "learned": {"overflow_rows": st["overflow_rows"], "predicted_rows": st["predicted_rows"],
                    "overflow_rate": round(st["overflow_rows"] / max(1, st["predicted_rows"]), 5),
                    "capacity_entries": st["capacity"], "heavy_rows": st["heavy_rows"],
                    "trained_weights": params is not None},
    }
    if "plain" in res:
        out["plain_gpu"] = {"ms_per_step": round(res["plain"], 4),
                            "value": round(flops / (res["plain"] * 1e-3) / 1e9, 3), "unit": "GFLOP/s",
                            "hbm_gbs": round(alg_bytes / (res["plain"] * 1e-3) / 1e9, 2)}
        out["speedup_vs_plain"] = round(res["plain"] / ms_step, 4)
    if "batched_exact" in res:
        out["batched_exact_gpu"] = {"ms_per_step": round(res["batched_exact"], 4),
                                    "value": round(flops / (res["batched_exact"] * 1e-3) / 1e9, 3)}
        ms_b, st_b = res["bound"]
        # what the trained network changes at this stage: the same learned pipeline with and without it
        out["predictor"] = {
            "trained_ms": round(ms_step, 4), "without_network_ms": round(ms_b, 4),
            "time_gain": round(ms_b / ms_step, 4),
            "slot_capacity_trained": st["capacity"], "slot_capacity_without_network": st_b["capacity"],
            "capacity_saved_entries": int(st_b["capacity"] - st["capacity"]),
            "rows_sized_by_network": st["predicted_rows"], "overflow_rows": st["overflow_rows"],
            "prediction": pred_stats,
            "note": "rows the network sizes are the heavy rows (more products than the sort bins take); they count "
                    "exactly before writing, so a short slot costs a bump allocation, not a redo"}
    if not args.no_cpu:
        t = time.time()
        cb, cbp = cpu_baseline_stage(pairs, args.cpu_products)
        out["cpu_baseline"] = cb
        out["cpu_baseline_parallel"] = cbp
        log(f"cpu baseline ({time.time() - t:.1f}s)")
    del A, B, pairs
    torch.cuda.empty_cache()
    if not args.headline_only and args.extra_stages:
        # the heavier stages of the same grid (k = 1, 2, 4: one high bit in the tile index, about nine times the
        # products of k = 3); k = 0 (about 3e10 output entries, 240 GB) does not fit one GPU next to its inputs
        for k in (int(x) for x in args.extra_stages.split(",")):
            out[f"stage_k{k}"] = bench_other_stage(args, h, tiles, k, params, margin, slack, tune)
    return out


def bench_other_stage(args, h, tiles, k, params, margin, slack, tune):
    """Another SUMMA stage of the same grid (learned against plain), e.g. the heavier k = 2: 12.5 G products."""
    import torch
    from bcl_amd.stage import LEARNED, PLAIN, StagePlan, stage_tiles
    A, B, pairs = stage_tiles(tiles, tiles, args.grid, k)
    steps, warm = max(2, args.steps // 2), 1
    t = time.time()
    plan = StagePlan(h, A, B, mode=LEARNED, params=params, margin=margin, slack=slack, tuning=tune)
    ms_l, _ = time_steps(plan.execute, steps, warm, h.stream)
    st = plan.stats()
    plan.close()
    torch.cuda.empty_cache()
    plain = StagePlan(h, A, B, mode=PLAIN, tuning=tune)
    ms_p, _ = time_steps(plain.execute, steps, warm, h.stream)
    st_p = plain.stats()
    plain.close()
    assert st_p["nnz"] == st["nnz"] and st_p["products"] == st["products"], (st_p, st)
    log(f"stage k={k}: learned {ms_l / steps:.2f} ms, plain {ms_p / steps:.2f} ms ({time.time() - t:.1f}s)")
    distinct_a = {id(x): x for x in A}
    distinct_b = {id(x): x for x in B}
    alg = sum(csr_bytes(x.m, x.nnz) for x in distinct_a.values()) + \
        sum(csr_bytes(x.m, x.nnz) for x in distinct_b.values()) + (st["rows"] + len(pairs)) * 8 + st["nnz"] * 8
    ms = ms_l / steps
    del A, B
    torch.cuda.empty_cache()
    return {"workload": f"rmat-s{args.scale}-ef16-grid{args.grid}-stage{k}", "steps": steps, "products": st["products"],
            "nnz_c": st["nnz"], "learned_ms": round(ms, 3), "value": round(2 * st["products"] / (ms * 1e-3) / 1e9, 3),
            "unit": "GFLOP/s", "plain_ms": round(ms_p / steps, 3), "speedup_vs_plain": round(ms_p / ms_l, 4),
            "hbm_gbs": round(alg / (ms * 1e-3) / 1e9, 2), "roofline_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "algorithmic_bytes": int(alg), "overflow_rows": st["overflow_rows"], "predicted_rows": st["predicted_rows"]}


def _ref_or_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    return O


def bench_config1(args):
    """Plain CSR SpGEMM C = A * A on one 16384^2 tile, exactly 8 uniform nnz/row (two-phase C ABI per product)."""
    import torch
    from bcl_amd import gen
    from bcl_amd.device import DeviceCSR, Handle, spgemm
    O = _ref_or_oracle()
    a = gen.uniform(16384, 16384, 8, seed=0xB0C1)
    h = Handle(0)
    A = DeviceCSR.from_host(a)
    C = spgemm(h, A, A)
    ms, per = time_steps(lambda: spgemm(h, A, A), args.steps, args.warmup, h.stream)
    ms_step = ms / args.steps
    products = int(np.diff(a.row_ptr)[a.col_idx].sum())
    nnz_c = C.nnz
    alg = 2 * csr_bytes(a.m, a.nnz) + (a.m + 1) * 8 + nnz_c * 8
    gbs = alg / (ms_step * 1e-3) / 1e9
    t = time.perf_counter()
    ref = O.ref_spgemm(a, a) if O.have_ref() else O.spgemm(a, a, dead_row_alloc=True)
    dt = time.perf_counter() - t
    assert ref.nnz == nnz_c
    return {"metric": "SpGEMM useful GFLOP/s, one 16384^2 tile, 8 nnz/row (two-phase, host-synchronised call)",
            "value": round(2 * products / (ms_step * 1e-3) / 1e9, 3), "unit": "GFLOP/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32 values, fp64 accumulate, int32 indices",
            "data": "synthetic",
            "config": {"workload": "uniform-16384-8perrow-AxA", "products": products, "nnz_c": nnz_c},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None, "algorithmic_bytes": alg,
                         "kernel": "whole two-phase call incl. host synchronisation (latency bound at this size)"},
            "cpu_baseline": {"value": round(2 * products / dt / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
                             "kind": "reference" if O.have_ref() else "port",
                             "sample": f"the full product, {dt:.2f} s serial BCL::gemm(CSRMatrix)"}}


def bench_spmm(args):
    """SpMM C += A * B: R-MAT scale 20 (ef 16) times dense 2^20 x 128 (N(0,1)), fp32 B and bf16 B."""
    import torch
    from bcl_amd import gen
    from bcl_amd.csr import CSR
    from bcl_amd.device import DeviceCSR, Handle, spmm
    O = _ref_or_oracle()
    a = gen.rmat(args.spmm_scale, ef=16, seed=args.seed)
    n = 128
    h = Handle(0)
    A = DeviceCSR.from_host(a)
    g = torch.Generator(device="cuda").manual_seed(7)
    B = torch.randn((a.n, n), device="cuda", dtype=torch.float32, generator=g)
    out = {}
    rows = np.random.default_rng(0).choice(a.m, 2048, replace=False)
    rows.sort()
    sub = CSR(len(rows), a.n, np.concatenate([[0], np.cumsum(np.diff(a.row_ptr)[rows])]).astype(np.int64),
              np.concatenate([a.col_idx[a.row_ptr[r]:a.row_ptr[r + 1]] for r in rows]),
              np.concatenate([a.vals[a.row_ptr[r]:a.row_ptr[r + 1]] for r in rows]))
    for name, Bt in (("fp32", B), ("bf16", B.to(torch.bfloat16))):
        C = torch.zeros((a.m, n), device="cuda", dtype=torch.float32)
        ms, _ = time_steps(lambda: spmm(h, A, Bt, C), args.steps, args.warmup, h.stream)
        ms_step = ms / args.steps
        C.zero_()
        spmm(h, A, Bt, C)
        torch.cuda.synchronize()
        bh = B.cpu().numpy()
        ref = O.spmm_f64(sub, bh)
        got = C.cpu().numpy()[rows]
        err = float(np.abs(got - ref).max() / np.abs(ref).max())
        absum = O.spmm_f64(CSR(sub.m, sub.n, sub.row_ptr, sub.col_idx, np.abs(sub.vals)), np.abs(bh))
        e_s = (np.abs(got - ref) / np.maximum(absum, 1e-300)).ravel()
        big = (np.abs(ref) >= 0.01 * absum).ravel()
        e_r = (np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)).ravel()[big]
        b_bytes = a.n * n * (4 if name == "fp32" else 2)
        alg = csr_bytes(a.m, a.nnz) + b_bytes + 2 * a.m * n * 4
        out[name] = {"ms": round(ms_step, 4), "gflops": round(2 * a.nnz * n / (ms_step * 1e-3) / 1e9, 2),
                     "hbm_gbs": round(alg / (ms_step * 1e-3) / 1e9, 2), "max_err_vs_fp64_rel_to_max": err,
                     "err_over_sum_abs_median": float(np.median(e_s)), "err_over_sum_abs_p99": float(np.quantile(e_s, 0.99)),
                     "rel_err_median": float(np.median(e_r)), "rel_err_p99": float(np.quantile(e_r, 0.99)),
                     "error_sample": "2048 random rows x 128 columns against the fp32-B product in FP64",
                     "algorithmic_bytes": alg}
    # reference CPU spmm_wrapper on a row sample
    rs = min(a.m, 65536)
    sub2 = CSR(rs, a.n, a.row_ptr[:rs + 1].copy(), a.col_idx[:a.row_ptr[rs]].copy(), a.vals[:a.row_ptr[rs]].copy())
    bh = B.cpu().numpy()
    cc = np.zeros((rs, n), np.float32)
    t = time.perf_counter()
    (O.ref_spmm if O.have_ref() else O.spmm)(sub2, bh, cc)
    dt = time.perf_counter() - t
    f = out["fp32"]
    return {"metric": "SpMM GFLOP/s + HBM GB/s %peak (C += A*B, R-MAT A, dense B n=128)", "value": f["gflops"],
            "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": f["ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (B fp32 or bf16)",
            "data": "synthetic", "config": {"workload": f"rmat-s{args.spmm_scale}-ef16-x-dense128", "nnz_a": a.nnz},
            "roofline": {"bound": "hbm", "achieved": f["hbm_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(f["hbm_gbs"] / HBM_PEAK_GBS, 5),
                         "traffic": profiled_traffic(f"spmm_rmat_s{args.spmm_scale}_n128"),
                         "algorithmic_bytes": f["algorithmic_bytes"], "kernel": "k_spmm (fp32 B)"},
            "fp32_b": out["fp32"], "bf16_b": out["bf16"],
            "cpu_baseline": {"value": round(2 * sub2.nnz * n / dt / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
                             "kind": "reference" if O.have_ref() else "port",
                             "sample": f"rows 0..{rs - 1} ({sub2.nnz} nnz), serial spmm_wrapper, {dt:.1f} s"}}


def bench_gemm(args):
    """Dense tile GEMM 8192^3, N(0,1): FP32 MFMA, BF16 MFMA, split-BF16 3-term; error vs FP64 for each."""
    import torch
    from bcl_amd.device import GEMM_BF16, GEMM_FP32, GEMM_SPLIT_BF16, Handle, gemm, gemm_f64_ref
    O = _ref_or_oracle()
    N = args.gemm_n
    h = Handle(0)
    g = torch.Generator(device="cuda").manual_seed(11)
    A = torch.randn((N, N), device="cuda", generator=g)
    B = torch.randn((N, N), device="cuda", generator=g)
    C64 = gemm_f64_ref(h, A, B)
    rows = np.arange(0, N, max(1, N // 16))[:16]
    ref_rows = O.gemm_f64(A[rows].cpu().numpy(), B.cpu().numpy())
    assert np.allclose(C64[rows].cpu().numpy(), ref_rows, rtol=1e-12, atol=1e-9), "FP64 GPU reference disagrees"
    peak = {GEMM_FP32: 157.3, GEMM_BF16: 2500.0, GEMM_SPLIT_BF16: 2500.0}
    issued = {GEMM_FP32: 1, GEMM_BF16: 1, GEMM_SPLIT_BF16: 3}
    res = {}
    S = gemm_f64_ref(h, A.abs(), B.abs())  # sum_k |a_ik| |b_kj|: the scale of each entry's rounding error
    pick = torch.randint(0, N * N, (1 << 20,), device="cuda", generator=torch.Generator(device="cuda").manual_seed(5))
    for mode, name in ((GEMM_FP32, "fp32"), (GEMM_BF16, "bf16"), (GEMM_SPLIT_BF16, "split_bf16_3")):
        C = torch.zeros((N, N), device="cuda")
        ms, _ = time_steps(lambda: gemm(h, A, B, C, mode=mode, beta=0.0), args.steps, args.warmup, h.stream)
        ms_step = ms / args.steps
        torch.cuda.synchronize()
        d = (C.double() - C64).abs()
        err = float((d.max() / C64.abs().max()).item())
        e_s = (d / S).flatten()[pick]
        big = (C64.abs() >= 0.01 * S).flatten()[pick]  # entries that are not cancellation residues
        e_r = ((d / C64.abs().clamp_min(1e-300)).flatten()[pick])[big]
        q = lambda x, p: float(torch.quantile(x, p).item())
        useful = 2.0 * N ** 3 / (ms_step * 1e-3) / 1e12
        res[name] = {"ms": round(ms_step, 3), "useful_tflops": round(useful, 2),
                     "mfma_util": round(useful * issued[mode] / peak[mode], 4),
                     "max_abs_err_rel_to_max_c": err,
                     "err_over_sum_abs_median": q(e_s, 0.5), "err_over_sum_abs_p99": q(e_s, 0.99),
                     "rel_err_median": q(e_r, 0.5), "rel_err_p99": q(e_r, 0.99),
                     "error_sample": "2^20 random entries of C; rel_err over entries with |C| >= 0.01 sum|a||b|"}
        del C, d
    # the fp64 product on the fp64 matrix cores (cblas_dgemm's place), on the same inputs widened exactly
    from bcl_amd.device import gemm_f64
    A64, B64 = A.double(), B.double()
    C = torch.zeros((N, N), device="cuda", dtype=torch.float64)
    ms, _ = time_steps(lambda: gemm_f64(h, A64, B64, C, beta=0.0), args.steps, args.warmup, h.stream)
    ms_step = ms / args.steps
    torch.cuda.synchronize()
    d = (C - C64).abs()
    useful = 2.0 * N ** 3 / (ms_step * 1e-3) / 1e12
    res["fp64"] = {"ms": round(ms_step, 3), "useful_tflops": round(useful, 2), "mfma_util": round(useful / 78.6, 4),
                   "peak_tflops": 78.6, "max_abs_err_rel_to_max_c": float((d.max() / C64.abs().max()).item()),
                   "err_over_sum_abs_max": float((d / S).max().item())}
    del A64, B64, C, d
    f = res["bf16"]
    # CPU: the oracle's float64 GEMM on 16 rows (labelled)
    t = time.perf_counter()
    O.gemm_f64(A[rows].cpu().numpy(), B.cpu().numpy(), nthreads=min(os.cpu_count() or 1, 16))
    dt = time.perf_counter() - t
    return {"metric": "Dense GEMM TFLOP/s and MFMA utilisation (8192^3)", "value": f["useful_tflops"],
            "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": f["ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 in, fp32 accumulate",
            "data": "synthetic", "config": {"workload": f"dense-gemm-{N}^3"},
            "roofline": {"bound": "mfma", "achieved": f["useful_tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": round(f["useful_tflops"] / 2500.0, 5), "traffic": None, "kernel": "k_gemm<1>"},
            "modes": res,
            "cpu_baseline": {"value": round(2 * len(rows) * N * N / dt / 1e9, 3), "unit": "GFLOP/s",
                             "cores": min(os.cpu_count() or 1, 16), "kind": "port",
                             "sample": f"{len(rows)} rows of C with the oracle's float64 blocked loop (no BLAS on the host)"}}


def bench_laplacian(args):
    """FP64 SpGEMM of the 2D 5-point Laplacian on a 4096 x 4096 grid: C = A * A."""
    from bcl_amd import gen
    from bcl_amd.csr import CSR
    from bcl_amd.device import DeviceCSR, Handle, spgemm
    O = _ref_or_oracle()
    g = args.lap_grid
    a = gen.laplacian2d(g)
    h = Handle(0)
    A = DeviceCSR.from_host(a)
    C = spgemm(h, A, A)
    ms, _ = time_steps(lambda: spgemm(h, A, A), args.steps, args.warmup, h.stream)
    ms_step = ms / args.steps
    products = int(np.diff(a.row_ptr)[a.col_idx].sum())
    # exact check on a band of rows against the oracle (integer values: exact in fp64)
    r0, r1 = g * (g // 2), g * (g // 2) + 4 * g
    c_rp = C.row_ptr[r0:r1 + 1].cpu().numpy()
    sub = CSR(r1 - r0, a.n, (a.row_ptr[r0:r1 + 1] - a.row_ptr[r0]).copy(), a.col_idx[a.row_ptr[r0]:a.row_ptr[r1]].copy(),
              a.vals[a.row_ptr[r0]:a.row_ptr[r1]].copy())
    ref = O.spgemm(sub, a)
    got_col = C.col_idx[int(c_rp[0]):int(c_rp[-1])].cpu().numpy()
    got_val = C.vals[int(c_rp[0]):int(c_rp[-1])].cpu().numpy()
    assert np.array_equal(got_col, ref.col_idx) and np.array_equal(got_val, ref.vals)
    alg = 2 * csr_bytes(a.m, a.nnz, 8, 8) + (a.m + 1) * 8 + C.nnz * 12
    gbs = alg / (ms_step * 1e-3) / 1e9
    t = time.perf_counter()
    sub2 = CSR(g * 64, a.n, a.row_ptr[:g * 64 + 1].copy(), a.col_idx[:a.row_ptr[g * 64]].copy(),
               a.vals[:a.row_ptr[g * 64]].copy())
    (O.ref_spgemm if O.have_ref() else O.spgemm)(sub2, a)
    dt = time.perf_counter() - t
    p2 = int(np.diff(a.row_ptr)[sub2.col_idx].sum())
    return {"metric": "FP64 SpGEMM GFLOP/s + HBM GB/s (2D 5-point Laplacian squared)",
            "value": round(2 * products / (ms_step * 1e-3) / 1e9, 3), "unit": "GFLOP/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"laplacian2d-{g}x{g}-squared", "nnz_a": a.nnz, "products": products,
                       "nnz_c": C.nnz, "nnz_c_expected": 218021892 if g == 4096 else None},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None, "algorithmic_bytes": alg,
                         "kernel": "whole two-phase call"},
            "cpu_baseline": {"value": round(2 * p2 / dt / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
                             "kind": "reference" if O.have_ref() else "port",
                             "sample": f"first {g * 64} rows, serial BCL::gemm(CSRMatrix<double>), {dt:.1f} s"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="stage", choices=["stage", "config1", "spmm", "gemm", "laplacian"])
    ap.add_argument("--spmm-scale", type=int, default=20)
    ap.add_argument("--gemm-n", type=int, default=8192)
    ap.add_argument("--lap-grid", type=int, default=4096)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--grid", type=int, default=8)
    ap.add_argument("--stage-k", type=int, default=3)
    ap.add_argument("--extra-stages", default="2,1,4",
                    help="other stages of the grid timed beside the headline (comma list; empty: none)")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-exact", action="store_true", help="skip the batched exact and without-network runs")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the stage workload (no extra stage, config 1, SpMM, GEMM or Laplacian keys)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-products", type=int, default=25_000_000)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build measures one GPU only: the stage's products run on one MI355X; "
                                   "multi-GPU SUMMA (tile broadcasts over RCCL/xGMI) is a design, not built"}))
        sys.exit(2)
    fns = {"stage": bench_stage, "config1": bench_config1, "spmm": bench_spmm, "gemm": bench_gemm,
           "laplacian": bench_laplacian}
    out = fns[args.workload](args)
    if args.workload == "stage" and not args.headline_only:
        # every other configuration of the benchmark, each with its own roofline and CPU baseline, as extra keys
        for name in ("config1", "spmm", "gemm", "laplacian"):
            t = time.time()
            sub = fns[name](args)
            for key in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
                sub.pop(key, None)
            out[name] = sub
            log(f"{name} done ({time.time() - t:.1f}s)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
