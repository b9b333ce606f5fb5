"""GPU kernels of the learned row-work predictor against the float64 CPU oracle.

- forward: |y_gpu - y_oracle| <= 1e-4 (1 + |y|)
- gradient of 0.5 * mean((y - t)^2): ||g_gpu - g_oracle|| / ||g_oracle|| <= 1e-4 and every parameter within
  1e-5 * max|g| + 1e-4 |g| (fp32 products, atomic fp32 sums over 64-sample partials)
- Adam step against a closed-form numpy update
- features: the documented formulas evaluated on the host for sampled rows
- training on a small dataset reduces the loss.
"""
import math

import numpy as np
import pytest
import torch

import oracle_py as O
from bcl_amd import gen, predictor
from bcl_amd.device import DeviceCSR

pytestmark = pytest.mark.gpu


def _rand_inputs(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1.5, size=(n, 16)).astype(np.float32)
    x[:, 15] = 1.0
    t = rng.uniform(0, 12, size=n).astype(np.float32)
    return x, t


def test_forward_matches_oracle(handle):
    x, _ = _rand_inputs(3000, 1)
    params = predictor.init_params(seed=3)
    y = predictor.forward(handle, torch.from_numpy(x).cuda(), torch.from_numpy(params).cuda()).cpu().numpy()
    y64, _, _ = O.mlp(x, params)
    assert np.all(np.abs(y - y64) <= 1e-4 * (1 + np.abs(y64)))


@pytest.mark.parametrize("n", [64, 1000, 4099])
def test_gradient_matches_oracle(handle, n):
    x, t = _rand_inputs(n, n)
    params = predictor.init_params(seed=n)
    g, loss = predictor.grad(handle, torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(),
                             torch.from_numpy(params).cuda())
    g = g.cpu().numpy().astype(np.float64)
    _, g64, l64 = O.mlp(x, params, t)
    assert abs(float(loss.item()) - l64) <= 1e-4 * abs(l64)
    rel = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    assert rel <= 1e-4, rel
    assert np.all(np.abs(g - g64) <= 1e-5 * np.abs(g64).max() + 1e-4 * np.abs(g64))


def test_gradient_detects_planted_error():
    """The oracle's gradient is itself pinned: it matches central finite differences of its own loss."""
    x, t = _rand_inputs(50, 7)
    params = predictor.init_params(seed=7).astype(np.float64)
    _, g64, _ = O.mlp(x, params.astype(np.float32), t)
    rng = np.random.default_rng(0)
    for idx in rng.choice(params.size, 40, replace=False):
        h = 1e-3
        pp = params.astype(np.float32).copy()
        pp[idx] += h
        _, _, lp = O.mlp(x, pp, t)
        pm = params.astype(np.float32).copy()
        pm[idx] -= h
        _, _, lm = O.mlp(x, pm, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g64[idx]) <= 2e-3 * max(1.0, abs(g64[idx])), (idx, fd, g64[idx])


def test_adam_step(handle):
    rng = np.random.default_rng(4)
    p = rng.standard_normal(5313).astype(np.float32)
    g = rng.standard_normal(5313).astype(np.float32)
    m = rng.standard_normal(5313).astype(np.float32) * 0.1
    v = np.abs(rng.standard_normal(5313).astype(np.float32)) * 0.1
    tp, tg, tm, tv = (torch.from_numpy(a.copy()).cuda() for a in (p, g, m, v))
    predictor.adam(handle, tp, tg, tm, tv, 3, 1e-2)
    b1, b2, eps = 0.9, 0.999, 1e-8
    m2 = b1 * m + (1 - b1) * g
    v2 = b2 * v + (1 - b2) * g * g
    p2 = p - 1e-2 * (m2 / (1 - b1 ** 3)) / (np.sqrt(v2 / (1 - b2 ** 3)) + eps)
    np.testing.assert_allclose(tp.cpu().numpy(), p2, rtol=1e-5, atol=1e-6)


def _features_host(a, b, i, prod):
    """Row features as documented in csrc/mlp.hpp, on the host."""
    cols = a.col_idx[a.row_ptr[i]:a.row_ptr[i + 1]]
    na = len(cols)
    step = na // 32 if na > 32 else 1
    samp = cols[::step][:32]
    lens = b.row_ptr[samp + 1] - b.row_ptr[samp]
    firsts = [b.col_idx[b.row_ptr[k]] for k, l in zip(samp, lens) if l > 0]
    lasts = [b.col_idx[b.row_ptr[k + 1] - 1] for k, l in zip(samp, lens) if l > 0]
    span = (max(lasts) - min(firsts) + 1) if firsts else 1.0
    lg = lambda v: math.log2(1 + max(v, 0))
    n = b.n
    avg_a = a.nnz / a.m
    avg_b = b.nnz / b.m
    e_span = span * (1 - math.exp(-prod / span))
    e_n = n * (1 - math.exp(-prod / n))
    lsum = lens.sum()
    f = [lg(na) / 16, lg(prod) / 32, lg(min(prod, n)) / 24, lg(lens.max() if len(lens) else 0) / 24,
         lg(lens.min() if len(lens) else 0) / 24, lg(n) / 24, lg(span) / 24,
         lg(sum(l_ - f_ + 1 for f_, l_ in zip(firsts, lasts))) / 32, lg(prod / span) / 16, lg(avg_b) / 16,
         lg(na / avg_a) / 8, lg(e_span) / 24, lg(e_n) / 24, (lens > 1).mean() if len(lens) else 0,
         lg(lsum * na / len(samp) if len(samp) else 0) / 32, 1.0]
    return np.array(f, np.float32)


def test_features(handle):
    a = gen.rmat(12, seed=3)
    A = DeviceCSR.from_host(a)
    feat, prod = predictor.features(handle, A, A)
    feat, prod = feat.cpu().numpy(), prod.cpu().numpy()
    _, prod_ref = O.row_nnz(a, a)
    np.testing.assert_array_equal(prod, prod_ref)
    rng = np.random.default_rng(0)
    rows = list(rng.choice(a.m, 200, replace=False)) + [int(np.argmax(a.row_nnz()))]
    for i in rows:
        np.testing.assert_allclose(feat[i], _features_host(a, a, i, int(prod[i])), rtol=1e-4, atol=1e-5)


def test_training_reduces_loss(handle):
    mats = predictor.training_matrices(small=True)
    x, t = predictor.build_dataset(handle, mats, max_rows_per_mat=20000)
    params, hist = predictor.train(handle, x, t, steps=400, batch=2048, log_every=100, verbose=False)
    assert hist[-1][1] < 0.2 * hist[0][1], hist
    ev = predictor.evaluate(handle, x, t, params)
    assert ev["median_rel_err"] < 0.5, ev
