"""Write the golden fixtures of the tile multiply from the reference's own CPU code.

Runs in the container that holds the reference sources: the reference's BCL::gemm(CSRMatrix...), spmm_wrapper,
SparseHashAccumulator::get_matrix and CSRMatrix::read_MatrixMarket are compiled from those sources into oracle/_ref (make -C oracle ref) and called
on small seeded inputs. Each fixture stores the inputs and the reference's outputs (.npz, no pickles):

  python tests/golden/make_golden.py

The inputs come from bcl_amd.gen (deterministic) and are stored in the fixture, so the checks do not depend on the
generators.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle_py as O  # noqa: E402
from bcl_amd import gen  # noqa: E402
from bcl_amd.csr import CSR  # noqa: E402


def pack(prefix, a: CSR):
    return {f"{prefix}_shape": np.array([a.m, a.n], np.int64), f"{prefix}_row_ptr": a.row_ptr.astype(np.int64),
            f"{prefix}_col_idx": a.col_idx.astype(np.int32), f"{prefix}_vals": a.vals}


def drop_rows(a: CSR, every: int) -> CSR:
    keep = np.ones(a.m, bool)
    keep[::every] = False
    rows = np.repeat(np.arange(a.m), a.row_nnz())
    mask = keep[rows]
    rp = np.zeros(a.m + 1, np.int64)
    rp[1:] = np.cumsum(np.bincount(rows[mask], minlength=a.m))
    return CSR(a.m, a.n, rp, a.col_idx[mask], a.vals[mask])


def spgemm_cases():
    cases = {}
    u = gen.uniform(256, 256, 8, seed=11)
    cases["spgemm_uniform256"] = (u, u, None)
    r = gen.rmat(9, seed=12)
    cases["spgemm_rmat9"] = (r, r, None)
    b1 = gen.banded(300, 200, bw=64, max_per_row=16, seed=13)
    b2 = gen.banded(200, 250, bw=64, max_per_row=16, seed=14)
    cases["spgemm_banded_rect"] = (drop_rows(b1, 5), b2, None)
    one = np.float32(1.0)
    a = CSR(1, 2, np.array([0, 2], np.int64), np.array([0, 1], np.int32), np.array([one, one]))
    b = CSR(2, 1, np.array([0, 1, 2], np.int64), np.array([0, 0], np.int32), np.array([one, -one]))
    cases["spgemm_zero_sum"] = (a, b, None)
    # integer values in {-2..2}: exact sums, cancellations keep explicit zeros
    rng = np.random.default_rng(15)
    ai = gen.uniform(64, 64, 6, seed=16)
    ai.vals = rng.integers(-2, 3, ai.nnz).astype(np.float32)
    cases["spgemm_int_cancel"] = (ai, ai, None)
    # C += A * B with a non-empty C (the pre-insert of spgemm.hpp:47-54)
    c0 = gen.uniform(256, 256, 4, seed=17)
    cases["spgemm_cplus"] = (u, u, c0)
    lap = gen.laplacian2d(12)
    cases["spgemm_laplacian12_f64"] = (lap, lap, None)
    return cases


def mm_cases():
    """Small MatrixMarket files (text) covering the reader's cases: comments, repeated entries (first wins),
    symmetric mirroring, pattern and integer fields, 0-based indices, empty rows."""
    rng = np.random.default_rng(40)

    def text(m, n, ents, field, sym, one):
        b = 1 if one else 0
        out = [f"%%MatrixMarket matrix coordinate {field} {sym}", "% written by make_golden.py", f"{m} {n} {len(ents)}"]
        for i, j, v in ents:
            out.append(f"{i + b} {j + b}" if field == "pattern" else
                       f"{i + b} {j + b} {int(v)}" if field == "integer" else f"{i + b} {j + b} {v:.9g}")
        return ("\n".join(out) + "\n").encode()

    def ents(m, n, count, lower):
        ii, jj = rng.integers(0, m, count), rng.integers(0, n, count)
        if lower:
            ii, jj = np.maximum(ii, jj), np.minimum(ii, jj)
        e = list(zip(ii.tolist(), jj.tolist(), rng.standard_normal(count).tolist()))
        e += [(e[q][0], e[q][1], float(rng.standard_normal())) for q in rng.integers(0, count, count // 5)]
        return e

    return {
        "general": (text(60, 40, ents(60, 40, 300, False), "real", "general", True), True),
        "symmetric": (text(50, 50, ents(50, 50, 250, True), "real", "symmetric", True), True),
        "pattern": (text(30, 70, ents(30, 70, 200, False), "pattern", "general", True), True),
        "integer": (text(40, 40, [(i, j, float(rng.integers(-9, 9))) for i, j, _ in ents(40, 40, 150, True)],
                         "integer", "symmetric", True), True),
        "zero_based": (text(25, 35, ents(25, 35, 120, False), "real", "general", False), False),
    }


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref/libbclref.so is missing: run `make -C oracle ref` next to the reference tree")
    out = {}
    for name, (a, b, c0) in spgemm_cases().items():
        c = O.ref_spgemm(a, b, c0)
        d = {**pack("a", a), **pack("b", b), **pack("c", c)}
        if c0 is not None:
            d.update(pack("c0", c0))
        out[name] = d
    # spmm_wrapper: C (m x n, ldc) += A (m x k) * B (k x n, ldb), with a non-zero initial C and padded ld
    a = gen.rmat(9, seed=21)
    rng = np.random.default_rng(22)
    n, ld = 24, 32
    bm = rng.standard_normal((a.n, ld)).astype(np.float32)
    cm = rng.standard_normal((a.m, ld)).astype(np.float32)
    c = cm.copy()
    O.ref_spmm(a, bm[:, :n].copy() if ld == n else bm, c) if ld == n else _ref_spmm_ld(a, bm, c, n)
    out["spmm_rmat9"] = {**pack("a", a), "b": bm, "c0": cm, "c": c, "n": np.array([n], np.int64)}
    # SparseHashAccumulator::get_matrix over 4 partial tiles
    tiles = [gen.banded(150, 120, bw=30, max_per_row=8, seed=30 + t) for t in range(4)]
    acc = O.ref_accumulate(tiles, 150, 120)
    d = {**pack("c", acc)}
    for t, tt in enumerate(tiles):
        d.update(pack(f"t{t}", tt))
    out["accumulate4"] = d
    # CSRMatrix::read_MatrixMarket on small files
    import tempfile
    mm = {}
    with tempfile.TemporaryDirectory() as tmp:
        for nm, (txt, one) in mm_cases().items():
            path = os.path.join(tmp, nm + ".mtx")
            with open(path, "wb") as f:
                f.write(txt)
            c = O.ref_mm_read(path, one_indexed=one)
            mm.update({f"{nm}_text": np.frombuffer(txt, np.uint8), f"{nm}_one": np.array([int(one)], np.int64),
                       **pack(nm, c)})
    out["mm_read"] = mm
    for name, d in out.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print(name, {k: v.shape for k, v in d.items() if k.endswith("col_idx")})


def _ref_spmm_ld(a, b, c, n):
    """spmm_wrapper with ldb = ldc = b.shape[1] > n (DMatrix tiles padded to the full tile width)."""
    import ctypes
    R = O.ref()
    arp = np.ascontiguousarray(a.row_ptr, np.int32)
    acol = np.ascontiguousarray(a.col_idx, np.int32)
    av = np.ascontiguousarray(a.vals, np.float32)
    P = ctypes.c_void_p
    R.ref_spmm_f32(a.m, n, a.n, a.nnz, c.ctypes.data_as(P), b.ctypes.data_as(P), av.ctypes.data_as(P),
                   arp.ctypes.data_as(P), acol.ctypes.data_as(P), b.shape[1], c.shape[1])


if __name__ == "__main__":
    main()
