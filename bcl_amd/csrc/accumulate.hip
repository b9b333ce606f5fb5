// Cross-stage accumulation (bclx_accumulate_tiles): C = sum of K partial m x n tiles, the GPU counterpart of
// SparseHashAccumulator::get_matrix (bcl/containers/sequential/SparseHashAccumulator.hpp:101-182) and of the
// pairwise sum_tiles_cusparse fold (bcl/containers/experimental/cuda/sequential/cusparse_util.cuh:251-265).
//
// The sum is computed as one SpGEMM: C = [T_1 T_2 ... T_K] * [I; I; ...; I], where the left operand stacks the tiles
// side by side (row i of it is the concatenation of row i of every tile, tile t's columns shifted by t * n) and the
// right operand stacks K identities. Every product is x * 1 (exact), so the two-phase SpGEMM returns exactly the
// union pattern, sorted, with each entry summed in float64 and rounded once.
#include <algorithm>

#include "spgemm.hpp"

namespace bclx {

template <typename T, typename RP>
struct TileIn {
  const RP* rp;
  const int* col;
  const T* val;
};

template <typename T, typename RP>
__global__ void k_stack_count(const TileIn<T, RP>* tiles, int K, int64_t m, int32_t* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  long long c = 0;
  for (int t = 0; t < K; t++) c += (long long)(tiles[t].rp[i + 1] - tiles[t].rp[i]);
  cnt[i] = (int32_t)c;
}

// one wavefront per row: copies row i of every tile into the stacked row, columns shifted by t * n
template <typename T, typename RP>
__global__ void k_stack_fill(const TileIn<T, RP>* tiles, int K, int64_t m, int64_t n, const int64_t* srp,
                             int* scol, T* sval) {
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= m) return;
  int64_t pos = srp[i];
  for (int t = 0; t < K; t++) {
    const RP a0 = tiles[t].rp[i], a1 = tiles[t].rp[i + 1];
    for (RP p = a0 + lane; p < a1; p += 64) {
      scol[pos + (p - a0)] = tiles[t].col[p] + (int)(t * n);
      sval[pos + (p - a0)] = tiles[t].val[p];
    }
    pos += (int64_t)(a1 - a0);
  }
}

template <typename T>
__global__ void k_ident_stack(int64_t rows, int64_t n, int64_t* rp, int* col, T* val) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r <= rows) rp[r] = r;
  if (r < rows) {
    col[r] = (int)(r % n);
    val[r] = (T)1;
  }
}

template <typename T, typename RP>
static void accumulate_impl(bclx_handle* h, int K, const bclx_csr* tiles, bclx_csr* C) {
  const int64_t m = tiles[0].m, n = tiles[0].n;
  hipStream_t s = h->stream;
  std::vector<TileIn<T, RP>> tin(K);
  for (int t = 0; t < K; t++)
    tin[t] = {static_cast<const RP*>(tiles[t].row_ptr), tiles[t].col_idx, static_cast<const T*>(tiles[t].vals)};
  int64_t total_in = 0;
  for (int t = 0; t < K; t++) total_in += tiles[t].nnz;
  BCLX_REQUIRE(K * n < INT32_MAX, BCLX_ERR_OVERFLOW, "K * n must fit int32 (stacked column index)");
  // workspace: tile table, counts, stacked A (int64 row_ptr), stacked identity B
  DevBuf tin_d(sizeof(TileIn<T, RP>) * K), cnt(sizeof(int32_t) * (m + 1)), srp(sizeof(int64_t) * (m + 1));
  DevBuf scol(sizeof(int) * std::max<int64_t>(total_in, 1)), sval(sizeof(T) * std::max<int64_t>(total_in, 1));
  const int64_t brows = (int64_t)K * n;
  DevBuf brp(sizeof(int64_t) * (brows + 1)), bcol(sizeof(int) * std::max<int64_t>(brows, 1)),
      bval(sizeof(T) * std::max<int64_t>(brows, 1));
  BCLX_HIP(hipMemcpyAsync(tin_d.p, tin.data(), sizeof(TileIn<T, RP>) * K, hipMemcpyHostToDevice, s));
  if (m > 0) {
    hipLaunchKernelGGL((k_stack_count<T, RP>), dim3(div_up(m, 256)), dim3(256), 0, s, tin_d.as<TileIn<T, RP>>(), K, m,
                       cnt.as<int32_t>());
    BCLX_HIP(hipGetLastError());
  }
  DevBuf part(sizeof(long long) * scan_partials(m));
  scan_counts(s, cnt.as<int32_t>(), m, srp.as<int64_t>(), part.as<long long>(), nullptr);
  if (m > 0) {
    hipLaunchKernelGGL((k_stack_fill<T, RP>), dim3(div_up(m, 4)), dim3(256), 0, s, tin_d.as<TileIn<T, RP>>(), K, m, n,
                       srp.as<int64_t>(), scol.as<int>(), sval.as<T>());
    BCLX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL((k_ident_stack<T>), dim3(div_up(brows + 1, 256)), dim3(256), 0, s, brows, n, brp.as<int64_t>(),
                     bcol.as<int>(), bval.as<T>());
  BCLX_HIP(hipGetLastError());
  bclx_csr A{m, brows, total_in, 8, tiles[0].dtype, srp.p, scol.as<int>(), sval.p};
  bclx_csr B{brows, n, brows, 8, tiles[0].dtype, brp.p, bcol.as<int>(), bval.p};
  TwoPhase<T, int64_t> tp(h, make_tile<T, int64_t>(&A, &B), BinTuning{});
  const int64_t nnz = tp.symbolic();
  if (C->col_idx == nullptr) {
    C->m = m;
    C->n = n;
    C->nnz = nnz;
    if (C->row_ptr) {
      BCLX_REQUIRE(C->ptr_width == 8 || nnz <= INT32_MAX, BCLX_ERR_OVERFLOW, "nnz does not fit an int32 row_ptr");
      const unsigned g = div_up(m + 1, 256);
      if (C->ptr_width == 8)
        hipLaunchKernelGGL(k_rowptr<int64_t>, dim3(g), dim3(256), 0, s, tp.F.R.off, (int64_t)0, m,
                           static_cast<int64_t*>(C->row_ptr));
      else
        hipLaunchKernelGGL(k_rowptr<int32_t>, dim3(g), dim3(256), 0, s, tp.F.R.off, (int64_t)0, m,
                           static_cast<int32_t*>(C->row_ptr));
      BCLX_HIP(hipGetLastError());
    }
  } else {
    BCLX_REQUIRE(C->row_ptr && C->vals, BCLX_ERR_INVALID, "C->row_ptr / C->vals are null");
    BCLX_REQUIRE(C->dtype == tiles[0].dtype, BCLX_ERR_INVALID, "C dtype must match the tiles");
    tp.numeric(C->col_idx, static_cast<T*>(C->vals), C->nnz, C->row_ptr, C->ptr_width);
    C->m = m;
    C->n = n;
    C->nnz = nnz;
  }
  BCLX_HIP(hipStreamSynchronize(s));  // the workspace above is released on return
}

}  // namespace bclx

using namespace bclx;

extern "C" {

bclx_status bclx_accumulate_tiles(bclx_handle* h, int ntiles, const bclx_csr* tiles, bclx_csr* C) {
  return guarded([&] {
    BCLX_REQUIRE(h && tiles && C && ntiles > 0, BCLX_ERR_INVALID, "null argument or ntiles <= 0");
    for (int t = 0; t < ntiles; t++) {
      const bclx_csr& x = tiles[t];
      BCLX_REQUIRE(x.m == tiles[0].m && x.n == tiles[0].n, BCLX_ERR_INVALID, "tiles must share their shape");
      BCLX_REQUIRE(x.dtype == tiles[0].dtype && x.ptr_width == tiles[0].ptr_width, BCLX_ERR_INVALID,
                   "tiles must share dtype and ptr_width");
      BCLX_REQUIRE(x.dtype == BCLX_F32 || x.dtype == BCLX_F64, BCLX_ERR_INVALID, "dtype must be fp32 or fp64");
      BCLX_REQUIRE(x.ptr_width == 4 || x.ptr_width == 8, BCLX_ERR_INVALID, "ptr_width must be 4 or 8");
      BCLX_REQUIRE(x.row_ptr && (x.nnz == 0 || (x.col_idx && x.vals)), BCLX_ERR_INVALID, "null tile arrays");
    }
    set_device(h);
    const bool f64 = tiles[0].dtype == BCLX_F64, w8 = tiles[0].ptr_width == 8;
    if (!f64 && !w8) accumulate_impl<float, int32_t>(h, ntiles, tiles, C);
    else if (!f64 && w8) accumulate_impl<float, int64_t>(h, ntiles, tiles, C);
    else if (f64 && !w8) accumulate_impl<double, int32_t>(h, ntiles, tiles, C);
    else accumulate_impl<double, int64_t>(h, ntiles, tiles, C);
  });
}

}  // extern "C"
