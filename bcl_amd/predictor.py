"""Learned row-work predictor for the batched SpGEMM stage.

A 16 -> 64 -> 64 -> 1 ReLU MLP maps per-row features of C = A * B (computed by bclx_row_features from A's row
and a sample of the B rows it selects) to log2(1 + nnz(C_i)). Inside the batched stage pipeline the prediction
sizes each row's output slot and hash table in place of the symbolic pass; rows whose true nnz exceeds the slot
are redone exactly. Training runs on the GPU through the library's own forward / backward / Adam kernels on
(features, exact symbolic nnz) pairs from seeded uniform, banded and R-MAT matrices.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import numpy as np
import torch

from ._lib import MLP_NFEAT, MLP_NPARAM, call
from .csr import CSR
from .device import DeviceCSR, Handle, spgemm

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS = os.path.join(DATA_DIR, "predictor_weights.npy")
WEIGHTS_META = os.path.join(DATA_DIR, "predictor_weights.json")
MIN_PRODUCTS = 32  # rows at or below this use the exact bound min(products, n) and are not predicted

H = 64
SHAPES = [("W1", (H, MLP_NFEAT)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H,)), ("b3", (1,))]


def _p(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if name.startswith("W"):
            fan_in = shp[-1] if len(shp) == 2 else H
            parts.append((rng.standard_normal(shp) * np.sqrt(2.0 / fan_in)).astype(np.float32).ravel())
        else:
            parts.append(np.zeros(shp, np.float32).ravel())
    p = np.concatenate(parts)
    assert p.size == MLP_NPARAM
    return p


def features(h: Handle, A: DeviceCSR, B: DeviceCSR):
    feat = torch.empty((A.m, MLP_NFEAT), dtype=torch.float32, device=A.vals.device)
    prod = torch.empty(A.m, dtype=torch.int64, device=A.vals.device)
    da, db = A.desc(), B.desc()
    call("bclx_row_features", h.h, ctypes.byref(da), ctypes.byref(db), _p(feat), _p(prod))
    return feat, prod


def forward(h: Handle, x: torch.Tensor, params: torch.Tensor) -> torch.Tensor:
    y = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
    call("bclx_mlp_forward", h.h, x.shape[0], _p(x), _p(params), _p(y))
    return y


def grad(h: Handle, x: torch.Tensor, t: torch.Tensor, params: torch.Tensor):
    g = torch.empty(MLP_NPARAM, dtype=torch.float32, device=x.device)
    loss = torch.empty(1, dtype=torch.float32, device=x.device)
    call("bclx_mlp_grad", h.h, x.shape[0], _p(x), _p(t), _p(params), _p(g), _p(loss))
    return g, loss


def adam(h: Handle, params, g, m, v, step: int, lr: float, b1=0.9, b2=0.999, eps=1e-8):
    call("bclx_mlp_adam", h.h, _p(params), _p(g), _p(m), _p(v), step, lr, b1, b2, eps)


def row_targets(h: Handle, A: DeviceCSR, B: DeviceCSR):
    """(features [m,16], products [m], exact nnz(C_i) [m]) from the symbolic pass of the two-phase path."""
    feat, prod = features(h, A, B)
    _, (prod_h, nnz_h) = spgemm(h, A, B, return_rows=True)
    return feat, prod, torch.from_numpy(nnz_h.astype(np.int64)).to(A.vals.device)


def training_matrices(small: bool = False):
    """Seeded training products: uniform, banded (bandwidth 64) and R-MAT scale 18..20 (whole and tiled)."""
    from . import gen
    mats = []
    n = 1 << (12 if small else 16)
    mats.append(("uniform8", gen.uniform(n, n, 8, seed=101), None))
    mats.append(("uniform3", gen.uniform(n, n, 3, seed=102), None))
    mats.append(("banded64", gen.banded(n, n, bw=64, max_per_row=16, seed=103), None))
    scales = [12, 13] if small else [18, 19, 20]
    for s in scales:
        a = gen.rmat(s, seed=200 + s)
        mats.append((f"rmat{s}", a, None))
        g = 4 if s >= 19 else 2
        tiles = gen.split_grid(a, g, g)
        for k in range(g):
            for (i, j) in [(0, 0), (g - 1, 0), (0, g - 1), (g // 2, g // 2)]:
                mats.append((f"rmat{s}_t{i}{k}{j}", tiles[(i, k)], tiles[(k, j)]))
    return mats


def build_dataset(h: Handle, mats, max_rows_per_mat: int = 400_000, seed: int = 0):
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    for name, a, b in mats:
        b = a if b is None else b
        A = DeviceCSR.from_host(a)
        B = A if b is a else DeviceCSR.from_host(b)
        feat, prod, nnz = row_targets(h, A, B)
        keep = (prod > MIN_PRODUCTS).nonzero().squeeze(1)
        if keep.numel() == 0:
            continue
        if keep.numel() > max_rows_per_mat:
            sel = torch.from_numpy(rng.choice(keep.numel(), max_rows_per_mat, replace=False)).to(keep.device)
            keep = keep[sel]
        xs.append(feat[keep])
        ts.append(torch.log2(1.0 + nnz[keep].to(torch.float32)))
    return torch.cat(xs), torch.cat(ts)


def train(h: Handle, x: torch.Tensor, t: torch.Tensor, steps: int = 3000, batch: int = 8192, lr: float = 3e-3,
          seed: int = 0, log_every: int = 500, verbose: bool = True):
    """Adam on 0.5 * mean((y - t)^2) with the library's gradient kernel. Returns (params numpy, loss history)."""
    dev = x.device
    params = torch.from_numpy(init_params(seed)).to(dev)
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    n = x.shape[0]
    t0 = time.time()
    for step in range(1, steps + 1):
        idx = torch.randint(0, n, (min(batch, n),), generator=gen).to(dev)
        xb, tb = x[idx].contiguous(), t[idx].contiguous()
        lr_t = lr * (0.5 * (1 + np.cos(np.pi * step / steps)) * 0.95 + 0.05)
        g, loss = grad(h, xb, tb, params)
        adam(h, params, g, m, v, step, float(lr_t))
        if step % log_every == 0 or step == 1:
            lv = float(loss.item())
            hist.append((step, lv))
            if verbose:
                print(f"[train] step {step} loss {lv:.5f} ({time.time() - t0:.1f}s)", flush=True)
    return params.cpu().numpy(), hist


def evaluate(h: Handle, x: torch.Tensor, t: torch.Tensor, params: np.ndarray, margin: float = 0.15, slack: int = 4):
    """Prediction quality on (features, log2(1+nnz)) pairs: overflow rate of the capacities the stage would use and
    the mean relative over-allocation."""
    p = torch.from_numpy(np.ascontiguousarray(params, np.float32)).to(x.device)
    y = forward(h, x, p)
    pred = torch.exp2(y.clamp(0, 40)) - 1
    true = torch.exp2(t) - 1
    cap = torch.ceil(pred * (1 + margin)) + slack
    overflow = (true > cap).float().mean().item()
    over = ((cap - true).clamp(min=0) / true.clamp(min=1)).mean().item()
    rel = ((pred - true).abs() / true.clamp(min=1)).median().item()
    return {"overflow_rate": overflow, "mean_over_allocation": over, "median_rel_err": rel}


def save(params: np.ndarray, meta: dict, out_dir: str = DATA_DIR):
    os.makedirs(out_dir, exist_ok=True)
    np.save(os.path.join(out_dir, os.path.basename(WEIGHTS)), np.ascontiguousarray(params, np.float32))
    with open(os.path.join(out_dir, os.path.basename(WEIGHTS_META)), "w") as f:
        json.dump(meta, f, indent=1)


def load() -> np.ndarray | None:
    if not os.path.exists(WEIGHTS):
        return None
    p = np.load(WEIGHTS, allow_pickle=False)
    assert p.dtype == np.float32 and p.size == MLP_NPARAM
    return p


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="train the row-work predictor on the GPU and save its weights")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--out", default=DATA_DIR, help="directory for predictor_weights.npy / .json")
    a = ap.parse_args()
    h = Handle(0)
    mats = training_matrices(small=a.small)
    x, t = build_dataset(h, mats)
    print(f"[train] {x.shape[0]} rows from {len(mats)} products", flush=True)
    perm = torch.randperm(x.shape[0], generator=torch.Generator().manual_seed(1)).to(x.device)
    nval = max(1, x.shape[0] // 10)
    xv, tv = x[perm[:nval]], t[perm[:nval]]
    xt, tt = x[perm[nval:]], t[perm[nval:]]
    params, hist = train(h, xt, tt, steps=a.steps)
    ev = evaluate(h, xv, tv, params)
    print("[train] validation", ev, flush=True)
    save(params, {"steps": a.steps, "rows": int(x.shape[0]), "history": hist, "validation": ev,
                  "matrices": [m[0] for m in mats]}, a.out)
