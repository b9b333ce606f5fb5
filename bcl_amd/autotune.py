"""Autotuner for the batched stage: picks the binning thresholds and the predicted-capacity margin.

Knobs (bclx_stage_opts.h_tuning, BinTuning in csrc/spgemm_kernels.hpp):
  work_per_lane  a row's group must have at least products / work_per_lane lanes (sets the group size per bin)
  heavy_need     rows needing more table entries than this take the heavy-row (rank-compact) accumulator
  load_shift     hash tables hold at least need << load_shift slots (1: load <= 1/2, 2: load <= 1/4)
and the learned capacity margin (capacity = ceil(pred * (1 + margin)) + slack).

Coordinate descent from the defaults: each knob in turn is swept with the others fixed, keeping the fastest setting.
Each candidate is a fresh learned plan on the real stage, timed over a few graph launches with HIP events. The
result is written to bcl_amd/data/autotune.json, which bench.py and StagePlan users load.

  python -m bcl_amd.autotune [--scale 22] [--stage-k 3] [--out DIR]
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT = {"work_per_lane": 64, "heavy_need": 4096, "load_shift": 1, "esc_min": 32, "esc_max": 4096, "bin_streams": 1,
           "margin": 0.15, "slack": 4}
SPACE = {"bin_streams": [1, 2, 4], "esc_max": [0, 512, 1024, 2048, 4096], "esc_min": [0, 16, 32, 64], "work_per_lane": [16, 32, 64, 128, 256],
         "heavy_need": [1024, 2048, 4096], "load_shift": [1, 2], "margin": [0.05, 0.1, 0.15, 0.25, 0.4]}
KNOBS = ("work_per_lane", "heavy_need", "load_shift", "esc_min", "esc_max", "bin_streams")


def load_tuning(path: str | None = None):
    """The shipped stage settings: (h_tuning array, margin, slack), or (None, defaults) when no table is saved."""
    path = path or os.path.join(DATA_DIR, "autotune.json")
    if not os.path.exists(path):
        return None, DEFAULT["margin"], DEFAULT["slack"]
    with open(path) as f:
        tj = json.load(f)
    return np.array(tj["tuning"], np.int32), float(tj.get("margin", DEFAULT["margin"])), int(tj.get("slack", DEFAULT["slack"]))


def time_plan(h, A, B, params, cfg, reps=3):
    import torch
    from .stage import LEARNED, StagePlan
    tun = np.array([cfg[k] for k in KNOBS], np.int32)
    plan = StagePlan(h, A, B, mode=LEARNED, params=params, margin=cfg["margin"], slack=cfg["slack"], tuning=tun)
    plan.execute()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for s, e in ev:
        s.record(h.stream)
        plan.execute()
        e.record(h.stream)
    torch.cuda.synchronize()
    ms = float(np.median([s.elapsed_time(e) for s, e in ev]))
    st = plan.stats()
    plan.close()
    return ms, st


def tune(h, A, B, params, log=print):
    cfg = dict(DEFAULT)
    best, st = time_plan(h, A, B, params, cfg)
    nnz = st["nnz"]
    log(f"[autotune] default {cfg}: {best:.2f} ms")
    trials = [(dict(cfg), best)]
    for knob, values in SPACE.items():
        for v in values:
            if v == cfg[knob]:
                continue
            cand = dict(cfg)
            cand[knob] = v
            ms, st = time_plan(h, A, B, params, cand)
            assert st["nnz"] == nnz, "a tuning choice changed the result"
            trials.append((cand, ms))
            log(f"[autotune] {knob}={v}: {ms:.2f} ms")
            if ms < best:
                best, cfg = ms, cand
    return cfg, best, trials


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--grid", type=int, default=8)
    ap.add_argument("--stage-k", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--out", default=DATA_DIR)
    a = ap.parse_args()
    from . import gen, predictor
    from .device import Handle
    from .stage import stage_tiles
    t0 = time.time()
    mat = gen.rmat(a.scale, seed=a.seed)
    tiles = gen.split_grid(mat, a.grid, a.grid)
    del mat
    h = Handle(0)
    A, B, _ = stage_tiles(tiles, tiles, a.grid, a.stage_k)
    params = predictor.load()
    cfg, best, trials = tune(h, A, B, params, log=lambda s: print(s, flush=True))
    os.makedirs(a.out, exist_ok=True)
    out = {"tuning": [cfg[k] for k in KNOBS], "margin": cfg["margin"],
           "slack": cfg["slack"], "ms": best, "default_ms": trials[0][1],
           "workload": f"rmat-s{a.scale}-ef16-grid{a.grid}-stage{a.stage_k}",
           "trials": [{"cfg": c, "ms": m} for c, m in trials], "seconds": time.time() - t0}
    with open(os.path.join(a.out, "autotune.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"[autotune] best {cfg}: {best:.2f} ms (default {trials[0][1]:.2f} ms)", flush=True)


if __name__ == "__main__":
    main()
