// Tile ingest and dense tile arithmetic on the GPU (bclx_csr_split_grid, bclx_dense_binary_op).
//
// bclx_csr_split_grid cuts one CSR matrix resident in HBM into the tiles of a grid, the way SPMatrix's constructor
// slices a matrix read from a file into its tiles (SPMatrix::init_, bcl/containers/SPMatrix.hpp:212-266, through
// CSRMatrix::get_slice_impl_, bcl/containers/sequential/CSRMatrix.hpp:238-276): tile (ti, tj) holds rows
// [ti * tile_m, (ti + 1) * tile_m) and columns [tj * tile_n, (tj + 1) * tile_n) of A, columns shifted to start at 0,
// entries in their order in A. With 288 GB of HBM a whole matrix fits next to its tiles, so the cut is two passes over
// A on the GPU instead of one host slice per tile:
//   1. count: one thread per (row, tile column) finds where the tile column's segment of the row starts (binary
//      search; A's columns are sorted within each row) and writes the segment length;
//   2. one scan per tile column over all rows turns the lengths into offsets; a tile's row_ptr is that scan
//      restricted to its rows, minus the value at its first row;
//   3. copy: one wavefront per row; each lane takes entries of the row, finds its tile column from the segment
//      starts staged in LDS, and writes the entry to its tile at (row offset + position in the segment).
// Bytes: A's row_ptr and col_idx are read twice (count, copy) and its values once; every tile is written once.
//
// bclx_dense_binary_op is DMatrix's elementwise arithmetic on a tile (DMatrix::binary_op / binary_op_inplace,
// bcl/containers/DMatrix.hpp:319-366, with operator+, -, *, / and += at :368-386): C = A op B per element, rows of
// ld-strided row-major tiles; C may be A (in place). HBM-bound streaming: 16-byte loads and stores where the rows
// are aligned, one pass.
#include <algorithm>
#include <vector>

#include "internal.hpp"
#include "spgemm.hpp"

namespace bclx {

constexpr int kSplitMaxGridN = 256;  // tile columns of one split (LDS of the copy kernel)

// cnt[tj * cstride + i] = entries of row i in tile column tj.
template <typename RP>
__global__ __launch_bounds__(256) void k_split_count(const RP* rp, const int32_t* col, int64_t m, int64_t tile_n,
                                                     int gn, int32_t* cnt, int64_t cstride) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * gn) return;
  const int64_t i = t / gn;
  const int tj = (int)(t - i * gn);
  const int64_t a0 = (int64_t)rp[i], a1 = (int64_t)rp[i + 1];
  // first position with col >= bound, for the tile column's two bounds
  auto lower = [&](int64_t bound) {
    int64_t lo = a0, hi = a1;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)col[mid] < bound) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  cnt[(int64_t)tj * cstride + i] = (int32_t)(lower((int64_t)(tj + 1) * tile_n) - lower((int64_t)tj * tile_n));
}

struct SplitTile {
  void* rp;      // tile rows + 1 entries of ptr_width
  int32_t* col;
  void* val;
  int64_t nnz;   // entries the tile holds (from the counting pass)
  int32_t ptr_width;
};

// One wavefront per row of A: row_ptr entries of its tiles, then its entries copied to their tiles. An entry that
// would land outside its tile (columns of A not sorted within the row, or not below n) is not written and sets *bad.
template <typename RP, typename T>
__global__ __launch_bounds__(256) void k_split_copy(const RP* rp, const int32_t* col, const T* val, int64_t m,
                                                    int64_t n, int64_t tile_m, int64_t tile_n, int gn,
                                                    const int64_t* off, int64_t ostride, const SplitTile* tiles,
                                                    int32_t* bad) {
  __shared__ int64_t s_seg[4][kSplitMaxGridN + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + w;
  const bool live = i < m;
  const int64_t a0 = live ? (int64_t)rp[i] : 0, a1 = live ? (int64_t)rp[i + 1] : 0;
  const int64_t ti = live ? i / tile_m : 0, li = i - ti * tile_m;
  const int64_t first = ti * tile_m;
  const int64_t rows = (m - first) < tile_m ? (m - first) : tile_m;
  // segment starts of the row's tile columns (positions in A), and the tiles' row_ptr entries of this row
  for (int tj = lane; live && tj <= gn; tj += 64) {
    int64_t lo = a0, hi = a1;
    const int64_t bound = (int64_t)tj * tile_n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)col[mid] < bound) lo = mid + 1;
      else hi = mid;
    }
    s_seg[w][tj] = tj == gn ? a1 : lo;
    if (tj < gn) {
      const int64_t* o = off + (int64_t)tj * ostride;
      const SplitTile& t = tiles[ti * gn + tj];
      const int64_t base = o[first];
      const int64_t v = o[i] - base;
      if (t.ptr_width == 8) static_cast<int64_t*>(t.rp)[li] = v;
      else static_cast<int32_t*>(t.rp)[li] = (int32_t)v;
      if (li == rows - 1) {  // the row that closes its tile row also writes the tile's last row_ptr entry
        const int64_t e = o[first + rows] - base;
        if (t.ptr_width == 8) static_cast<int64_t*>(t.rp)[rows] = e;
        else static_cast<int32_t*>(t.rp)[rows] = (int32_t)e;
      }
    }
  }
  __syncthreads();
  for (int64_t p = a0 + lane; p < a1; p += 64) {
    const int64_t c = col[p];
    const int64_t tj = c / tile_n;
    if (c < 0 || c >= n || tj >= gn) {
      *bad = 1;
      continue;
    }
    const SplitTile& t = tiles[ti * gn + tj];
    const int64_t* o = off + tj * ostride;
    const int64_t dst = (o[i] - o[first]) + (p - s_seg[w][tj]);
    if (p < s_seg[w][tj] || p >= s_seg[w][tj + 1] || dst < 0 || dst >= t.nnz) {
      *bad = 1;
      continue;
    }
    t.col[dst] = (int32_t)(c - tj * tile_n);
    static_cast<T*>(t.val)[dst] = val[p];
  }
}

// tot[ti * gn + tj] = entries of tile (ti, tj).
__global__ void k_split_totals(const int64_t* off, int64_t ostride, int64_t m, int64_t tile_m, int gm, int gn,
                               int64_t* tot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= gm * gn) return;
  const int ti = t / gn, tj = t - ti * gn;
  const int64_t r0 = (int64_t)ti * tile_m, r1 = (r0 + tile_m) < m ? (r0 + tile_m) : m;
  const int64_t* o = off + (int64_t)tj * ostride;
  tot[t] = o[r1] - o[r0];
}

template <typename RP, typename T>
void split_grid(bclx_handle* h, const bclx_csr* A, int64_t tile_m, int64_t tile_n, int gm, int gn, bclx_csr* tiles) {
  hipStream_t s = h->stream;
  const int64_t m = A->m;
  const int64_t cstride = (m + 3) / 4 * 4;      // 16-byte aligned count rows
  const int64_t ostride = (m + 1 + 3) / 4 * 4;  // 16-byte aligned offset rows
  DevBuf cnt(sizeof(int32_t) * (size_t)(cstride * gn));
  DevBuf off(sizeof(int64_t) * (size_t)(ostride * gn));
  DevBuf partial(sizeof(long long) * (size_t)scan_partials(m));
  DevBuf tot(sizeof(int64_t) * (size_t)(gm * gn));
  const RP* rp = static_cast<const RP*>(A->row_ptr);
  if (m > 0) {
    const int64_t nt = m * gn;
    hipLaunchKernelGGL((k_split_count<RP>), dim3(div_up(nt, 256)), dim3(256), 0, s, rp, A->col_idx, m, tile_n, gn,
                       cnt.as<int32_t>(), cstride);
    BCLX_HIP(hipGetLastError());
  }
  for (int tj = 0; tj < gn; tj++)
    scan_counts(s, cnt.as<int32_t>() + (int64_t)tj * cstride, m, off.as<int64_t>() + (int64_t)tj * ostride,
                partial.as<long long>(), nullptr);
  hipLaunchKernelGGL(k_split_totals, dim3(div_up((int64_t)gm * gn, 256)), dim3(256), 0, s, off.as<int64_t>(), ostride,
                     m, tile_m, gm, gn, tot.as<int64_t>());
  BCLX_HIP(hipGetLastError());
  std::vector<int64_t> nnz((size_t)gm * gn);
  BCLX_HIP(hipMemcpyAsync(nnz.data(), tot.p, sizeof(int64_t) * nnz.size(), hipMemcpyDeviceToHost, s));
  BCLX_HIP(hipStreamSynchronize(s));
  bool sizes_only = true;
  for (int t = 0; t < gm * gn; t++) sizes_only = sizes_only && tiles[t].row_ptr == nullptr;
  for (int t = 0; t < gm * gn; t++) {
    const int ti = t / gn, tj = t % gn;
    const int64_t r0 = (int64_t)ti * tile_m, c0 = (int64_t)tj * tile_n;
    bclx_csr& d = tiles[t];
    d.m = std::min<int64_t>(tile_m, m - r0);
    d.n = std::min<int64_t>(tile_n, A->n - c0);
    d.dtype = A->dtype;
    if (sizes_only) {
      d.nnz = nnz[t];
      continue;
    }
    BCLX_REQUIRE(d.nnz == nnz[t], BCLX_ERR_INVALID,
                 "tile " + std::to_string(t) + ": nnz " + std::to_string(d.nnz) + " given, " + std::to_string(nnz[t]) +
                     " needed (sizes come from a call with every row_ptr NULL)");
    BCLX_REQUIRE(d.ptr_width == 4 || d.ptr_width == 8, BCLX_ERR_INVALID, "tile ptr_width must be 4 or 8");
    BCLX_REQUIRE(d.ptr_width == 8 || nnz[t] <= INT32_MAX, BCLX_ERR_OVERFLOW,
                 "tile " + std::to_string(t) + " has more than 2^31 - 1 entries: use ptr_width 8");
    BCLX_REQUIRE(d.row_ptr != nullptr && (nnz[t] == 0 || (d.col_idx != nullptr && d.vals != nullptr)),
                 BCLX_ERR_INVALID, "null tile array");
  }
  if (sizes_only) return;
  std::vector<SplitTile> st((size_t)gm * gn);
  for (int t = 0; t < gm * gn; t++)
    st[t] = SplitTile{tiles[t].row_ptr, tiles[t].col_idx, tiles[t].vals, nnz[t], tiles[t].ptr_width};
  DevBuf dt(sizeof(SplitTile) * st.size() + sizeof(int32_t));
  int32_t* bad = reinterpret_cast<int32_t*>(dt.as<char>() + sizeof(SplitTile) * st.size());
  BCLX_HIP(hipMemcpyAsync(dt.p, st.data(), sizeof(SplitTile) * st.size(), hipMemcpyHostToDevice, s));
  BCLX_HIP(hipMemsetAsync(bad, 0, sizeof(int32_t), s));
  if (m > 0) {
    hipLaunchKernelGGL((k_split_copy<RP, T>), dim3(div_up(m, 4)), dim3(256), 0, s, rp, A->col_idx,
                       static_cast<const T*>(A->vals), m, A->n, tile_m, tile_n, gn, off.as<int64_t>(), ostride,
                       dt.as<SplitTile>(), bad);
    BCLX_HIP(hipGetLastError());
  } else {  // one row of tiles with no rows: each row_ptr is {0}
    for (int t = 0; t < gm * gn; t++) BCLX_HIP(hipMemsetAsync(tiles[t].row_ptr, 0, (size_t)tiles[t].ptr_width, s));
  }
  int32_t h_bad = 0;
  BCLX_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  BCLX_HIP(hipStreamSynchronize(s));  // the scratch buffers are released on return
  BCLX_REQUIRE(h_bad == 0, BCLX_ERR_INVALID, "A's columns must be sorted within each row and below A->n");
}

// ---- dense elementwise ---------------------------------------------------------------------------------------

template <int OP, typename T>
__device__ inline T apply_op(T a, T b) {
  if (OP == 0) return a + b;
  if (OP == 1) return a - b;
  if (OP == 2) return a * b;
  return a / b;
}

// Rows of VEC-element vectors: block (x, y) covers vectors [x * 256, x * 256 + 256) of rows y, y + gridDim.y, ...
template <int OP, typename T, int VEC>
__global__ __launch_bounds__(256) void k_binary_op(int64_t m, int64_t nvec, const T* A, int64_t lda, const T* B,
                                                   int64_t ldb, T* C, int64_t ldc) {
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nvec) return;
  for (int64_t i = blockIdx.y; i < m; i += gridDim.y) {
    const vec_t a = *reinterpret_cast<const vec_t*>(A + i * lda + v * VEC);
    const vec_t b = *reinterpret_cast<const vec_t*>(B + i * ldb + v * VEC);
    vec_t c;
#pragma unroll
    for (int q = 0; q < VEC; q++) c[q] = apply_op<OP, T>(a[q], b[q]);
    *reinterpret_cast<vec_t*>(C + i * ldc + v * VEC) = c;
  }
}

template <int OP, typename T>
void binary_op(hipStream_t s, int64_t m, int64_t n, const T* A, int64_t lda, const T* B, int64_t ldb, T* C,
               int64_t ldc) {
  constexpr int V = 16 / sizeof(T);
  auto aligned = [](const void* p, int64_t ld) {
    return (reinterpret_cast<uintptr_t>(p) % 16) == 0 && (ld * (int64_t)sizeof(T)) % 16 == 0;
  };
  // a full tile (ld == n for all three) is one long row
  if (lda == n && ldb == n && ldc == n) {
    n *= m;
    lda = ldb = ldc = n;
    m = 1;
  }
  const bool vec = n % V == 0 && aligned(A, lda) && aligned(B, ldb) && aligned(C, ldc);
  const int64_t nvec = vec ? n / V : n;
  const unsigned gx = div_up(nvec, 256);
  const unsigned gy = (unsigned)std::max<int64_t>(1, std::min<int64_t>(m, std::max<int64_t>(1, 4096 / (int64_t)gx)));
  if (vec)
    hipLaunchKernelGGL((k_binary_op<OP, T, V>), dim3(gx, gy), dim3(256), 0, s, m, nvec, A, lda, B, ldb, C, ldc);
  else
    hipLaunchKernelGGL((k_binary_op<OP, T, 1>), dim3(gx, gy), dim3(256), 0, s, m, nvec, A, lda, B, ldb, C, ldc);
  BCLX_HIP(hipGetLastError());
}

template <typename T>
void binary_op_t(hipStream_t s, int op, int64_t m, int64_t n, const void* A, int64_t lda, const void* B, int64_t ldb,
                 void* C, int64_t ldc) {
  const T* a = static_cast<const T*>(A);
  const T* b = static_cast<const T*>(B);
  T* c = static_cast<T*>(C);
  switch (op) {
    case 0: binary_op<0, T>(s, m, n, a, lda, b, ldb, c, ldc); break;
    case 1: binary_op<1, T>(s, m, n, a, lda, b, ldb, c, ldc); break;
    case 2: binary_op<2, T>(s, m, n, a, lda, b, ldb, c, ldc); break;
    default: binary_op<3, T>(s, m, n, a, lda, b, ldb, c, ldc); break;
  }
}

}  // namespace bclx

using namespace bclx;

extern "C" {

bclx_status bclx_csr_split_grid(bclx_handle* h, const bclx_csr* A, int64_t tile_m, int64_t tile_n, int32_t ntiles,
                                bclx_csr* tiles) {
  return guarded([&] {
    BCLX_REQUIRE(h && A && tiles, BCLX_ERR_INVALID, "null argument");
    BCLX_REQUIRE(tile_m > 0 && tile_n > 0, BCLX_ERR_INVALID, "tile sizes must be positive");
    BCLX_REQUIRE(A->m >= 0 && A->n >= 0 && A->nnz >= 0, BCLX_ERR_INVALID, "bad matrix shape");
    BCLX_REQUIRE(A->ptr_width == 4 || A->ptr_width == 8, BCLX_ERR_INVALID, "A->ptr_width must be 4 or 8");
    BCLX_REQUIRE(A->dtype == BCLX_F32 || A->dtype == BCLX_F64, BCLX_ERR_INVALID, "A->dtype must be fp32 or fp64");
    BCLX_REQUIRE(A->row_ptr && (A->nnz == 0 || (A->col_idx && A->vals)), BCLX_ERR_INVALID, "null array in A");
    const int64_t gm = std::max<int64_t>(1, (A->m + tile_m - 1) / tile_m);
    const int64_t gn = std::max<int64_t>(1, (A->n + tile_n - 1) / tile_n);
    BCLX_REQUIRE(gn <= kSplitMaxGridN, BCLX_ERR_UNSUPPORTED,
                 "more than " + std::to_string(kSplitMaxGridN) + " tile columns in one split");
    BCLX_REQUIRE((int64_t)ntiles == gm * gn, BCLX_ERR_INVALID,
                 "tiles must hold ceil(m / tile_m) * ceil(n / tile_n) = " + std::to_string(gm * gn) + " descriptors");
    set_device(h);
    const bool w8 = A->ptr_width == 8, f64 = A->dtype == BCLX_F64;
    if (!w8 && !f64) split_grid<int32_t, float>(h, A, tile_m, tile_n, (int)gm, (int)gn, tiles);
    else if (!w8) split_grid<int32_t, double>(h, A, tile_m, tile_n, (int)gm, (int)gn, tiles);
    else if (!f64) split_grid<int64_t, float>(h, A, tile_m, tile_n, (int)gm, (int)gn, tiles);
    else split_grid<int64_t, double>(h, A, tile_m, tile_n, (int)gm, (int)gn, tiles);
  });
}

bclx_status bclx_dense_binary_op(bclx_handle* h, int32_t op, int32_t dtype, int64_t m, int64_t n, const void* d_A,
                                 int64_t lda, const void* d_B, int64_t ldb, void* d_C, int64_t ldc) {
  return guarded([&] {
    BCLX_REQUIRE(h != nullptr, BCLX_ERR_INVALID, "null handle");
    BCLX_REQUIRE(op >= 0 && op <= 3, BCLX_ERR_INVALID, "op must be 0 (+), 1 (-), 2 (*) or 3 (/)");
    BCLX_REQUIRE(dtype == BCLX_F32 || dtype == BCLX_F64, BCLX_ERR_INVALID, "dtype must be fp32 or fp64");
    BCLX_REQUIRE(m >= 0 && n >= 0 && lda >= n && ldb >= n && ldc >= n, BCLX_ERR_INVALID, "bad shape or ld");
    if (m == 0 || n == 0) return;
    BCLX_REQUIRE(d_A && d_B && d_C, BCLX_ERR_INVALID, "null array");
    set_device(h);
    if (dtype == BCLX_F32) binary_op_t<float>(h->stream, op, m, n, d_A, lda, d_B, ldb, d_C, ldc);
    else binary_op_t<double>(h->stream, op, m, n, d_A, lda, d_B, ldb, d_C, ldc);
  });
}

}  // extern "C"
