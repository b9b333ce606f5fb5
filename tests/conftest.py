import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbclx.so on cuda:0)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def handle():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bcl_amd.device import Handle
    h = Handle(0)
    yield h
    h.close()


def spgemm_value_bound(a, b, c64_vals, absum_vals, dtype):
    """Per-entry bound for a result accumulated in float64 and rounded once to `dtype`:
    |c - c64| <= u * |c64| + 1e-13 * sum|a||b| (u = unit roundoff of dtype; the second term absorbs
    float64 reordering)."""
    u = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -52
    return u * np.abs(c64_vals) + 1e-13 * absum_vals
