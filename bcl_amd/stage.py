"""One SUMMA stage of BCL::gemm(SPMatrix, SPMatrix, SPMatrix) on one GPU: the grid x grid local products
A_ik * B_kj of stage k (bcl/containers/algorithms/spgemm.hpp:181-242), run through bclx_stage_* as

  mode 0  plain: per-product two-phase launches with host synchronisation (the per-tile library call)
  mode 2  batched exact: one graph-captured pipeline over all products, symbolic + numeric
  mode 1  batched learned: one graph-captured pipeline, MLP-predicted capacities replace the symbolic pass
"""
from __future__ import annotations

import ctypes

import numpy as np
from ._lib import bclx_csr, bclx_stage_opts, call
from .csr import CSR
from .device import DeviceCSR, Handle

PLAIN, LEARNED, BATCHED_EXACT = 0, 1, 2
# bclx_stage_opts.h_tuning defaults: work_per_lane, heavy_need, load_shift, esc_min, esc_max, bin_streams
TUNING_DEFAULTS = (64, 4096, 1, 32, 4096, 1)


def full_tuning(t) -> np.ndarray:
    """A complete BCLX_TUNING_LEN table: the given leading knobs, defaults for the rest."""
    t = [int(x) for x in t]
    return np.array(t + list(TUNING_DEFAULTS[len(t):]), np.int32)
STAT_NAMES = ["products", "nnz", "overflow_rows", "predicted_rows", "capacity", "rows", "heavy_rows", "sorted_rows"]


class StagePlan:
    def __init__(self, h: Handle, A: list[DeviceCSR], B: list[DeviceCSR], mode: int = LEARNED,
                 params: np.ndarray | None = None, margin: float = 0.15, slack: int = 4, use_graph: bool = True,
                 tuning: np.ndarray | None = None):
        assert len(A) == len(B) and len(A) > 0
        self.h = h
        self.A, self.B = A, B  # keep the inputs alive as long as the plan
        self.npairs = len(A)
        da = (bclx_csr * self.npairs)(*[a.desc() for a in A])
        db = (bclx_csr * self.npairs)(*[b.desc() for b in B])
        o = bclx_stage_opts()
        o.mode = mode
        o.use_graph = int(use_graph)
        o.margin = margin
        o.slack = slack
        self._params = None if params is None else np.ascontiguousarray(params, np.float32)
        o.h_mlp_params = self._params.ctypes.data_as(ctypes.c_void_p) if self._params is not None else None
        o.persistent_blocks = 0
        self._tuning = None if tuning is None else np.ascontiguousarray(full_tuning(tuning), np.int32)
        o.h_tuning = self._tuning.ctypes.data_as(ctypes.c_void_p) if self._tuning is not None else None
        p = ctypes.c_void_p()
        call("bclx_stage_plan_create", h.h, self.npairs, da, db, ctypes.byref(o), ctypes.byref(p))
        self.p = p
        self.mode = mode

    def execute(self):
        call("bclx_stage_execute", self.p)

    def stats(self) -> dict:
        out = np.zeros(len(STAT_NAMES), np.int64)
        call("bclx_stage_stats", self.p, out.ctypes.data_as(ctypes.c_void_p), len(STAT_NAMES))
        return dict(zip(STAT_NAMES, [int(x) for x in out]))

    def predictions(self) -> dict:
        """The last execution's network predictions: arrays product, row, pred (float), cap, nnz per sized row."""
        cnt = ctypes.c_int64()
        call("bclx_stage_predictions", self.p, 0, None, None, None, None, None, ctypes.byref(cnt))
        k = cnt.value
        out = {"product": np.zeros(k, np.int32), "row": np.zeros(k, np.int32), "pred": np.zeros(k, np.float32),
               "cap": np.zeros(k, np.int32), "nnz": np.zeros(k, np.int32)}
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        call("bclx_stage_predictions", self.p, k, ptr(out["product"]), ptr(out["row"]), ptr(out["pred"]),
             ptr(out["cap"]), ptr(out["nnz"]), ctypes.byref(cnt))
        return out

    def result_desc(self, p: int) -> bclx_csr:
        d = bclx_csr()
        call("bclx_stage_result", self.p, p, ctypes.byref(d))
        return d

    def result_host(self, p: int) -> CSR:
        d = self.result_desc(p)
        dt = np.float32 if d.dtype == 0 else np.float64
        m, nnz = d.m, d.nnz
        rp = np.zeros(m + 1, np.int64)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz, dt)
        d2h(rp, d.row_ptr)
        if nnz:
            d2h(col, d.col_idx)
            d2h(val, d.vals)
        return CSR(m, d.n, rp, col, val)

    def close(self):
        if self.p:
            call("bclx_stage_plan_destroy", self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hiplib = None


def d2h(dst: np.ndarray, src_ptr: int) -> None:
    """Copy dst.nbytes from a raw device pointer returned by the library into a host array."""
    global _hiplib
    if _hiplib is None:
        _hiplib = ctypes.CDLL("libamdhip64.so")
        _hiplib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hiplib.hipMemcpy.restype = ctypes.c_int
    rc = _hiplib.hipMemcpy(dst.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src_ptr), dst.nbytes, 2)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy device-to-host failed ({rc})")


def stage_tiles(tiles_a: dict, tiles_b: dict, grid: int, k: int, device="cuda"):
    """Device tiles of SUMMA stage k: lists A_p = A[i, k], B_p = B[k, j] for p = i * grid + j."""
    dev_a = {i: DeviceCSR.from_host(tiles_a[(i, k)], device=device) for i in range(grid)}
    dev_b = {j: DeviceCSR.from_host(tiles_b[(k, j)], device=device) for j in range(grid)}
    A = [dev_a[i] for i in range(grid) for j in range(grid)]
    B = [dev_b[j] for i in range(grid) for j in range(grid)]
    host_pairs = [(tiles_a[(i, k)], tiles_b[(k, j)]) for i in range(grid) for j in range(grid)]
    return A, B, host_pairs
