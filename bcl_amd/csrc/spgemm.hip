// SpGEMM host orchestration: the per-product two-phase path (bclx_spgemm_symbolic / bclx_spgemm_numeric) and the
// flat-row workspace shared with the batched stage pipeline (stage.hip).
#include "spgemm.hpp"

#include <algorithm>
#include <cstring>

namespace bclx {

// ---------------------------------------------------------------------------------------------------------------
// small fill kernels (used instead of hipMemsetAsync so every pipeline can be captured in a HIP graph)
// ---------------------------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

void fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return;
  unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_i32, dim3(g), dim3(256), 0, s, p, n, v);
  BCLX_HIP(hipGetLastError());
}
void fill_u64(hipStream_t s, unsigned long long* p, int64_t n, unsigned long long v) {
  if (n <= 0) return;
  unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u64, dim3(g), dim3(256), 0, s, p, n, v);
  BCLX_HIP(hipGetLastError());
}

void ensure_dense_slots(bclx_handle* h, int nslots, size_t ndoubles) {
  if (h->dense_nslots >= nslots && h->dense_slot_doubles >= ndoubles) return;
  int ns = std::max(nslots, h->dense_nslots);
  size_t nd = std::max(ndoubles, h->dense_slot_doubles);
  BCLX_HIP(hipStreamSynchronize(h->stream));
  h->dense_slots.alloc((size_t)ns * nd * sizeof(double));
  int64_t tot = (int64_t)ns * (int64_t)nd;
  unsigned g = (unsigned)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fill_f64, dim3(g), dim3(256), 0, h->stream, h->dense_slots.as<double>(), tot);
  BCLX_HIP(hipGetLastError());
  BCLX_HIP(hipStreamSynchronize(h->stream));
  h->dense_nslots = ns;
  h->dense_slot_doubles = nd;
}

// ---------------------------------------------------------------------------------------------------------------
// FlatRows
// ---------------------------------------------------------------------------------------------------------------

template <typename T, typename RP>
void FlatRows<T, RP>::setup(bclx_handle* h_, const std::vector<TileDev<T, RP>>& tiles_in, bool with_cap) {
  h = h_;
  tiles = tiles_in;
  ntiles = (int)tiles.size();
  nflat = 0;
  max_n = 1;
  std::vector<uint16_t> blk;
  for (int t = 0; t < ntiles; t++) {
    tiles[t].flat0 = nflat;
    int64_t pm = ((int64_t)tiles[t].m + kRowBlock - 1) / kRowBlock * kRowBlock;
    if (pm == 0) pm = kRowBlock;  // every product owns at least one block
    for (int64_t b = 0; b < pm / kRowBlock; b++) blk.push_back((uint16_t)t);
    nflat += pm;
    max_n = std::max<int64_t>(max_n, tiles[t].n);
  }
  BCLX_REQUIRE(ntiles <= 65535, BCLX_ERR_INVALID, "at most 65535 products per batch");
  BCLX_REQUIRE(nflat < INT32_MAX, BCLX_ERR_OVERFLOW, "too many rows in one batch (flat rows must fit int32)");
  nblocks = nflat / kRowBlock;
  Arena& A = arena;
  A.used = 0;
  size_t o_tiles = A.reserve(sizeof(TileDev<T, RP>) * ntiles);
  size_t o_blk = A.reserve(sizeof(uint16_t) * nblocks);
  size_t o_prod = A.reserve(sizeof(int64_t) * nflat);
  size_t o_nnz = A.reserve(sizeof(int32_t) * nflat);
  size_t o_cap = A.reserve(sizeof(int32_t) * (with_cap ? nflat : 1));
  size_t o_off = A.reserve(sizeof(int64_t) * (nflat + 1));
  size_t o_bc = A.reserve(sizeof(int32_t) * kNumBins * kCountStride);
  size_t o_bi = A.reserve(sizeof(int32_t) * kNumBins * nflat);
  size_t o_oc = A.reserve(sizeof(int32_t) * 3);
  size_t o_oi = A.reserve(sizeof(int32_t) * (with_cap ? nflat : 1));
  size_t o_pi = A.reserve(sizeof(int32_t) * (with_cap ? nflat : 1));
  size_t o_cur = A.reserve(sizeof(unsigned long long) * 2);
  size_t o_tmp = A.reserve(sizeof(int32_t) * nflat);
  size_t o_stats = A.reserve(sizeof(unsigned long long) * 8);
  size_t nsb = (size_t)((nflat + kScanTile - 1) / kScanTile) + 1;
  size_t o_part = A.reserve(sizeof(long long) * nsb);
  size_t o_tot = A.reserve(sizeof(int64_t) * 2);
  A.commit();
  R.tiles = A.at<TileDev<T, RP>>(o_tiles);
  R.blk_tile = A.at<uint16_t>(o_blk);
  R.prod = A.at<int64_t>(o_prod);
  R.nnz = A.at<int32_t>(o_nnz);
  R.cap = A.at<int32_t>(o_cap);
  R.off = A.at<int64_t>(o_off);
  R.bin_count = A.at<int32_t>(o_bc);
  R.bin_items = A.at<int32_t>(o_bi);
  R.bin_stride = nflat;
  R.ovf_count = A.at<int32_t>(o_oc);
  R.ovf_items = A.at<int32_t>(o_oi);
  R.pred_count = R.ovf_count + 1;
  R.long_count = R.ovf_count + 2;
  R.pred_items = A.at<int32_t>(o_pi);
  R.cursor = A.at<unsigned long long>(o_cur);
  R.heavy_queue = R.cursor + 1;
  R.tmp_items = A.at<int32_t>(o_tmp);
  R.stats = A.at<unsigned long long>(o_stats);
  scan_partial = A.at<long long>(o_part);
  scan_total = A.at<int64_t>(o_tot);
  BCLX_HIP(hipMemcpyAsync(const_cast<TileDev<T, RP>*>(R.tiles), tiles.data(), sizeof(TileDev<T, RP>) * ntiles,
                          hipMemcpyHostToDevice, h->stream));
  BCLX_HIP(hipMemcpyAsync(const_cast<uint16_t*>(R.blk_tile), blk.data(), sizeof(uint16_t) * nblocks,
                          hipMemcpyHostToDevice, h->stream));
  BCLX_HIP(hipStreamSynchronize(h->stream));
}

template <typename T, typename RP>
void FlatRows<T, RP>::upload_tiles() {
  BCLX_HIP(hipMemcpyAsync(const_cast<TileDev<T, RP>*>(R.tiles), tiles.data(), sizeof(TileDev<T, RP>) * ntiles,
                          hipMemcpyHostToDevice, h->stream));
}

template <typename T, typename RP>
void FlatRows<T, RP>::reset_counters(hipStream_t s) {
  fill_i32(s, R.bin_count, kNumBins * kCountStride, 0);
  fill_i32(s, R.ovf_count, 3, 0);
  fill_u64(s, R.stats, 8, 0);
}

template <typename T, typename RP>
void FlatRows<T, RP>::analyze(hipStream_t s, bool do_bin, long long pred_min, bool with_stats) {
  const unsigned grid = (unsigned)((nblocks + kAnaSub - 1) / kAnaSub);
  fill_i32(s, R.long_count, 1, 0);
  hipLaunchKernelGGL((k_analyze<T, RP>), dim3(grid), dim3(kRowBlock), 0, s, R, (int64_t)nblocks, do_bin ? 1 : 0,
                     pred_min, with_stats ? 1 : 0);
  BCLX_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_analyze_long<T, RP>), dim3((unsigned)h->num_cus * 8), dim3(kRowBlock), 0, s, R,
                     do_bin ? 1 : 0, pred_min, with_stats ? 1 : 0);
  BCLX_HIP(hipGetLastError());
}

template <typename T, typename RP>
void FlatRows<T, RP>::bin(hipStream_t s, int src) {
  fill_i32(s, R.bin_count, kNumBins * kCountStride, 0);
  hipLaunchKernelGGL((k_bin<T, RP>), dim3((unsigned)nblocks), dim3(kRowBlock), 0, s, R, src);
  BCLX_HIP(hipGetLastError());
}

template <typename T, typename RP>
void FlatRows<T, RP>::scan(hipStream_t s, const int32_t* cnt, int64_t* off, int64_t off_add) {
  const int64_t nb = (nflat + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanBlock), 0, s, cnt, nflat, scan_partial);
  BCLX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, s, scan_partial, nb, scan_total);
  BCLX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(kScanBlock), 0, s, cnt, nflat, scan_partial, off, off_add);
  BCLX_HIP(hipGetLastError());
}

template <int B, int MODE, typename T, typename RP>
static void launch_bin_kernel(const RowsDev<T, RP>& R, hipStream_t s, unsigned grid) {
  constexpr int TS = bin_table(B), G = bin_group(B);
  constexpr int BLOCK = G < 256 ? 256 : G;
  hipLaunchKernelGGL((k_hash<TS, G, BLOCK, MODE, T, RP>), dim3(grid), dim3(BLOCK), 0, s, R, B);
  BCLX_HIP(hipGetLastError());
}

template <int MODE, typename T, typename RP>
static void launch_lds_bin(const RowsDev<T, RP>& R, hipStream_t s, int b, unsigned grid) {
  switch (b) {
    case 1: launch_bin_kernel<1, MODE>(R, s, grid); break;
    case 2: launch_bin_kernel<2, MODE>(R, s, grid); break;
    case 3: launch_bin_kernel<3, MODE>(R, s, grid); break;
    case 4: launch_bin_kernel<4, MODE>(R, s, grid); break;
    case 5: launch_bin_kernel<5, MODE>(R, s, grid); break;
    case 6: launch_bin_kernel<6, MODE>(R, s, grid); break;
    case 7: launch_bin_kernel<7, MODE>(R, s, grid); break;
    case 8: launch_bin_kernel<8, MODE>(R, s, grid); break;
    case 9: launch_bin_kernel<9, MODE>(R, s, grid); break;
    default: throw Error{BCLX_ERR_INVALID, "bad LDS bin"};
  }
}

int64_t scan_partials(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void scan_counts(hipStream_t s, const int32_t* cnt, int64_t n, int64_t* off, long long* partial, int64_t* total) {
  if (n == 0) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned long long*>(off), (int64_t)1, 0ull);
    if (total) hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned long long*>(total), (int64_t)1, 0ull);
    BCLX_HIP(hipGetLastError());
    return;
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanBlock), 0, s, cnt, n, partial);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, s, partial, nb, total);
  hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(kScanBlock), 0, s, cnt, n, partial, off, (int64_t)0);
  BCLX_HIP(hipGetLastError());
}

int rows_per_block(int b) {
  const int G = bin_group(b);
  const int BLOCK = G < 256 ? 256 : G;
  return BLOCK / G;
}

template <int K, typename T, typename RP>
void launch_esc(const RowsDev<T, RP>& R, hipStream_t s, int mode, int b, unsigned grid) {
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_esc<K, 0, T, RP>), dim3(grid), dim3(64 * kEscWaves), 0, s, R, b); break;
    case 1: hipLaunchKernelGGL((k_esc<K, 1, T, RP>), dim3(grid), dim3(64 * kEscWaves), 0, s, R, b); break;
    case 2: hipLaunchKernelGGL((k_esc<K, 2, T, RP>), dim3(grid), dim3(64 * kEscWaves), 0, s, R, b); break;
    default: hipLaunchKernelGGL((k_esc<K, 3, T, RP>), dim3(grid), dim3(64 * kEscWaves), 0, s, R, b); break;
  }
  BCLX_HIP(hipGetLastError());
}

template <int W, typename T, typename RP>
void launch_esc_block(const RowsDev<T, RP>& R, hipStream_t s, int mode, int b, unsigned grid) {
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_esc_block<W, 0, T, RP>), dim3(grid), dim3(64 * W), 0, s, R, b); break;
    case 1: hipLaunchKernelGGL((k_esc_block<W, 1, T, RP>), dim3(grid), dim3(64 * W), 0, s, R, b); break;
    case 2: hipLaunchKernelGGL((k_esc_block<W, 2, T, RP>), dim3(grid), dim3(64 * W), 0, s, R, b); break;
    default: hipLaunchKernelGGL((k_esc_block<W, 3, T, RP>), dim3(grid), dim3(64 * W), 0, s, R, b); break;
  }
  BCLX_HIP(hipGetLastError());
}

template <typename T, typename RP>
void FlatRows<T, RP>::run_bin(hipStream_t s, int mode, int b, int64_t count_hint) {
  // count_hint >= 0: exact row count known on the host (grid sized to it); < 0: persistent grid.
  if (b == 0 || count_hint == 0) return;
  // rows of at most 32 products (bin 18) and of 33-64 (ESC bin 11): sorted in registers by groups of G lanes
  auto launch_small = [&](auto g_tag, auto n_tag) {
    constexpr int G = decltype(g_tag)::value, N = decltype(n_tag)::value;
    unsigned grid = count_hint >= 0 ? (unsigned)std::max<int64_t>(1, div_up(count_hint, (64 / G) * kSmallWaves))
                                    : (unsigned)(h->num_cus * 8);
    grid = std::min<unsigned>(grid, 1u << 20);
#define BCLX_SMALL(M, ONE) \
  hipLaunchKernelGGL((k_esc_small<M, ONE, G, N, T, RP>), dim3(grid), dim3(64 * kSmallWaves), 0, s, R, b)
    if (ntiles == 1) {
      switch (mode) { case 0: BCLX_SMALL(0, true); break; case 1: BCLX_SMALL(1, true); break;
                      case 2: BCLX_SMALL(2, true); break; default: BCLX_SMALL(3, true); break; }
    } else {
      switch (mode) { case 0: BCLX_SMALL(0, false); break; case 1: BCLX_SMALL(1, false); break;
                      case 2: BCLX_SMALL(2, false); break; default: BCLX_SMALL(3, false); break; }
    }
#undef BCLX_SMALL
    BCLX_HIP(hipGetLastError());
  };
  if (b == kSmallBin) {
    launch_small(std::integral_constant<int, kSmallGroup>{}, std::integral_constant<int, kSmallMax>{});
    return;
  }
  if (b == kEscBin0) {
    launch_small(std::integral_constant<int, 16>{}, std::integral_constant<int, 64>{});
    return;
  }
  if (b >= kEscBin0 && b < kEscBin0 + kEscWaveBins) {
    unsigned grid = count_hint >= 0 ? (unsigned)std::max<int64_t>(1, div_up(count_hint, kEscWaves))
                                    : (unsigned)(h->num_cus * 8);
    grid = std::min<unsigned>(grid, 1u << 20);
    switch (b - kEscBin0) {
      case 1: launch_esc<2>(R, s, mode, b, grid); break;
      case 2: launch_esc<4>(R, s, mode, b, grid); break;
      default: launch_esc<8>(R, s, mode, b, grid); break;  // bin 14 (bin 11 took the register sort above)
    }
    return;
  }
  if (b >= kEscBin0 + kEscWaveBins) {
    unsigned grid = count_hint >= 0 ? (unsigned)std::max<int64_t>(1, count_hint) : (unsigned)(h->num_cus * 8);
    grid = std::min<unsigned>(grid, 1u << 20);
    switch ((64 << (b - kEscBin0)) / kEscChunk) {  // wavefronts per row
      case 2: launch_esc_block<2>(R, s, mode, b, grid); break;
      case 4: launch_esc_block<4>(R, s, mode, b, grid); break;
      default: launch_esc_block<8>(R, s, mode, b, grid); break;
    }
    return;
  }
  if (b == kHeavyBin && (max_n + 31) / 32 <= kHRWords) {
    const int64_t resident = (mode == 0 ? 2 : 1) * (int64_t)h->num_cus;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(count_hint >= 0 ? count_hint : resident, resident));
    fill_u64(s, R.heavy_queue, 1, 0ull);
    hipLaunchKernelGGL((k_order_heavy<T, RP>), dim3(1), dim3(1024), 0, s, R);
    switch (mode) {
      case 0: hipLaunchKernelGGL((k_heavy_rank<0, T, RP>), dim3(grid), dim3(hr_block<0>()), 0, s, R, kHeavyBin); break;
      case 1: hipLaunchKernelGGL((k_heavy_rank<1, T, RP>), dim3(grid), dim3(hr_block<1>()), 0, s, R, kHeavyBin); break;
      case 2: hipLaunchKernelGGL((k_heavy_rank<2, T, RP>), dim3(grid), dim3(hr_block<2>()), 0, s, R, kHeavyBin); break;
      default: hipLaunchKernelGGL((k_heavy_rank<3, T, RP>), dim3(grid), dim3(hr_block<3>()), 0, s, R, kHeavyBin); break;
    }
    BCLX_HIP(hipGetLastError());
    return;
  }
  if (b == kHeavyBin) {  // global dense accumulator (tiles too wide for a window index)
    // prepare_heavy() made the accumulators and bitmaps for every resident workgroup before any graph capture;
    // nothing is allocated or attributed here
    const int64_t words = (max_n + 31) / 32;
    const bool lds_bitmap = heavy_lds_bitmap();
    BCLX_REQUIRE(h->dense_nslots >= h->num_cus && h->dense_slot_doubles >= (size_t)max_n &&
                     (lds_bitmap || gbits.bytes >= sizeof(unsigned) * (size_t)words * h->num_cus),
                 BCLX_ERR_INVALID, "heavy-row accumulators were not prepared");
    int nslots = (int)std::min<int64_t>(count_hint >= 0 ? count_hint : h->num_cus, (int64_t)h->num_cus);
    if (nslots < 1) nslots = 1;
    const size_t shm = lds_bitmap ? (size_t)words * 4 : 0;
    double* dense = h->dense_slots.as<double>();
    const int64_t sd = (int64_t)h->dense_slot_doubles;
    unsigned* gb = lds_bitmap ? nullptr : gbits.as<unsigned>();
#define BCLX_HEAVY(M, LB)                                                                                 \
  do {                                                                                                    \
    hipLaunchKernelGGL((k_heavy<M, LB, T, RP>), dim3(nslots), dim3(kHeavyThreads), shm, s, R, kHeavyBin, dense, sd, \
                       gb, words);                                                                        \
  } while (0)
    if (lds_bitmap) {
      switch (mode) { case 0: BCLX_HEAVY(0, true); break; case 1: BCLX_HEAVY(1, true); break;
                      case 2: BCLX_HEAVY(2, true); break; default: BCLX_HEAVY(3, true); break; }
    } else {
      switch (mode) { case 0: BCLX_HEAVY(0, false); break; case 1: BCLX_HEAVY(1, false); break;
                      case 2: BCLX_HEAVY(2, false); break; default: BCLX_HEAVY(3, false); break; }
    }
#undef BCLX_HEAVY
    BCLX_HIP(hipGetLastError());
    return;
  }
  const int rpb = rows_per_block(b);
  const int64_t per_cu = 8;  // resident blocks per CU targeted by the persistent grid
  unsigned grid;
  if (count_hint >= 0) grid = (unsigned)std::max<int64_t>(1, (count_hint + rpb - 1) / rpb);
  else grid = (unsigned)(h->num_cus * per_cu);
  grid = std::min<unsigned>(grid, 1u << 20);
  switch (mode) {
    case 0: launch_lds_bin<0>(R, s, b, grid); break;
    case 1: launch_lds_bin<1>(R, s, b, grid); break;
    case 2: launch_lds_bin<2>(R, s, b, grid); break;
    default: launch_lds_bin<3>(R, s, b, grid); break;
  }
}

template <typename T, typename RP>
void FlatRows<T, RP>::compact(hipStream_t s, const int64_t* dst_off, int* dst_col, T* dst_val, const int* src_col,
                              const T* src_val, int64_t max_out, int32_t* chunk_row) {
  hipLaunchKernelGGL(k_chunk_rows, dim3((unsigned)nblocks), dim3(kRowBlock), 0, s, dst_off, nflat, (long long)max_out,
                     chunk_row);
  BCLX_HIP(hipGetLastError());
  const int64_t nchunks = std::max<int64_t>(1, div_up(max_out, kCompactChunk));
  BCLX_REQUIRE(nchunks < INT32_MAX, BCLX_ERR_OVERFLOW, "output too large for one compaction launch");
  hipLaunchKernelGGL((k_compact<T, RP>), dim3((unsigned)nchunks), dim3(kRowBlock), 0, s, R, dst_off, nflat, chunk_row,
                     (long long)max_out, dst_col, dst_val, src_col, src_val);
  BCLX_HIP(hipGetLastError());
}

template <typename T, typename RP>
void FlatRows<T, RP>::read_counts(hipStream_t s, int32_t* dst) {
  int32_t raw[kNumBins * kCountStride];
  BCLX_HIP(hipMemcpyAsync(raw, R.bin_count, sizeof(raw), hipMemcpyDeviceToHost, s));
  BCLX_HIP(hipStreamSynchronize(s));
  for (int b = 0; b < kNumBins; b++) dst[b] = raw[b * kCountStride];
}

template <typename T, typename RP>
int64_t FlatRows<T, RP>::read_stat(int q) {
  unsigned long long v = 0;
  BCLX_HIP(hipMemcpyAsync(&v, R.stats + q, sizeof(v), hipMemcpyDeviceToHost, h->stream));
  BCLX_HIP(hipStreamSynchronize(h->stream));
  return (int64_t)v;
}

template <typename T, typename RP>
bool FlatRows<T, RP>::heavy_lds_bitmap() const {
  return ((max_n + 31) / 32) * 4 <= (int64_t)h->lds_per_block - 4096;
}

template <typename T, typename RP>
void FlatRows<T, RP>::prepare_heavy() {
  // The rank-slice kernel keeps everything in LDS. The global dense fallback (tiles wider than 2^19 columns) needs
  // one accumulator of max_n doubles per resident workgroup, and a zeroed bitmap per workgroup when the tile is too
  // wide for an LDS bitmap; the kernel leaves both zeroed behind every row.
  if ((max_n + 31) / 32 <= kHRWords) return;
  ensure_dense_slots(h, h->num_cus, (size_t)max_n);
  const int64_t words = (max_n + 31) / 32;
  if (heavy_lds_bitmap()) {
    const int shm = (int)(words * 4);
    BCLX_HIP(hipFuncSetAttribute((const void*)k_heavy<0, true, T, RP>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
    BCLX_HIP(hipFuncSetAttribute((const void*)k_heavy<1, true, T, RP>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
    BCLX_HIP(hipFuncSetAttribute((const void*)k_heavy<2, true, T, RP>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
    BCLX_HIP(hipFuncSetAttribute((const void*)k_heavy<3, true, T, RP>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
    return;
  }
  const size_t need = sizeof(unsigned) * (size_t)words * (size_t)h->num_cus;
  if (gbits.bytes < need) {
    gbits.alloc(need);
    fill_i32(h->stream, gbits.as<int32_t>(), (int64_t)(need / 4), 0);
    BCLX_HIP(hipStreamSynchronize(h->stream));
  }
}

template struct FlatRows<float, int32_t>;
template struct FlatRows<float, int64_t>;
template struct FlatRows<double, int32_t>;
template struct FlatRows<double, int64_t>;

// ---------------------------------------------------------------------------------------------------------------
// two-phase per-product path
// ---------------------------------------------------------------------------------------------------------------

static void check_csr(const bclx_csr* x, const char* name) {
  BCLX_REQUIRE(x != nullptr, BCLX_ERR_INVALID, std::string(name) + " is null");
  BCLX_REQUIRE(x->m >= 0 && x->n >= 0 && x->nnz >= 0, BCLX_ERR_INVALID, std::string(name) + ": negative size");
  BCLX_REQUIRE(x->ptr_width == 4 || x->ptr_width == 8, BCLX_ERR_INVALID, std::string(name) + ": ptr_width must be 4 or 8");
  BCLX_REQUIRE(x->dtype == BCLX_F32 || x->dtype == BCLX_F64, BCLX_ERR_INVALID,
               std::string(name) + ": dtype must be BCLX_F32 or BCLX_F64");
  BCLX_REQUIRE(x->row_ptr != nullptr, BCLX_ERR_INVALID, std::string(name) + ": row_ptr is null");
  BCLX_REQUIRE(x->nnz == 0 || (x->col_idx && x->vals), BCLX_ERR_INVALID, std::string(name) + ": null arrays");
  BCLX_REQUIRE(x->m < INT32_MAX && x->n < INT32_MAX, BCLX_ERR_INVALID, std::string(name) + ": dimension >= 2^31");
  BCLX_REQUIRE(x->ptr_width == 8 || x->nnz <= INT32_MAX, BCLX_ERR_OVERFLOW, std::string(name) + ": nnz > int32 row_ptr");
}

void check_pair(const bclx_csr* A, const bclx_csr* B) {
  check_csr(A, "A");
  check_csr(B, "B");
  BCLX_REQUIRE(A->n == B->m, BCLX_ERR_INVALID,
               "inner dimensions differ: A is " + std::to_string(A->m) + "x" + std::to_string(A->n) + ", B is " +
                   std::to_string(B->m) + "x" + std::to_string(B->n));
  BCLX_REQUIRE(A->dtype == B->dtype, BCLX_ERR_INVALID, "A and B value types differ");
  BCLX_REQUIRE(A->ptr_width == B->ptr_width, BCLX_ERR_INVALID, "A and B row_ptr widths differ");
}

template <typename T, typename RP>
TileDev<T, RP> make_tile(const bclx_csr* A, const bclx_csr* B) {
  TileDev<T, RP> t{};
  t.a_rp = static_cast<const RP*>(A->row_ptr);
  t.a_col = A->col_idx;
  t.a_val = static_cast<const T*>(A->vals);
  t.b_rp = static_cast<const RP*>(B->row_ptr);
  t.b_col = B->col_idx;
  t.b_val = static_cast<const T*>(B->vals);
  t.m = (int32_t)A->m;
  t.n = (int32_t)B->n;
  t.krows = B->m;
  t.avg_a = A->m ? (float)((double)A->nnz / (double)A->m) : 0.f;
  t.avg_b = B->m ? (float)((double)B->nnz / (double)B->m) : 0.f;
  return t;
}
template TileDev<float, int32_t> make_tile<float, int32_t>(const bclx_csr*, const bclx_csr*);
template TileDev<float, int64_t> make_tile<float, int64_t>(const bclx_csr*, const bclx_csr*);
template TileDev<double, int32_t> make_tile<double, int32_t>(const bclx_csr*, const bclx_csr*);
template TileDev<double, int64_t> make_tile<double, int64_t>(const bclx_csr*, const bclx_csr*);

template <typename T, typename RP>
int64_t TwoPhase<T, RP>::symbolic() {
  hipStream_t s = h->stream;
  F.reset_counters(s);
  F.analyze(s, true);
  int32_t* counts = sym_counts;
  F.read_counts(s, counts);
  if (counts[kHeavyBin] > 0) F.prepare_heavy();
  reuse_bins = true;  // the hash bins and the heavy bin are chosen by the row's size bound, which the count replaces
  for (int b = 1; b <= kHeavyBin; b++) reuse_bins = reuse_bins && counts[b] == 0;
  for (int b = 1; b < kNumBins; b++) F.run_bin(s, 0, b, counts[b]);
  F.scan(s, F.R.nnz, F.R.off, 0);
  int64_t tot = 0;
  BCLX_HIP(hipMemcpyAsync(&tot, F.scan_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  BCLX_HIP(hipStreamSynchronize(s));
  nnz_c = tot;
  return tot;
}

template <typename T, typename RP>
void TwoPhase<T, RP>::numeric(int* c_col, T* c_val, int64_t c_cap, void* row_ptr, int ptr_width) {
  BCLX_REQUIRE(nnz_c >= 0, BCLX_ERR_INVALID, "numeric phase before the symbolic phase");
  BCLX_REQUIRE(c_cap >= nnz_c, BCLX_ERR_INVALID, "C arrays smaller than nnz from the symbolic phase");
  BCLX_REQUIRE(ptr_width == 8 || nnz_c <= INT32_MAX, BCLX_ERR_OVERFLOW,
               "nnz(C) = " + std::to_string(nnz_c) + " does not fit an int32 row_ptr; request ptr_width 8");
  hipStream_t s = h->stream;
  if (F.tiles[0].c_col != c_col || F.tiles[0].c_val != c_val) {
    F.tiles[0].c_col = c_col;
    F.tiles[0].c_val = c_val;
    F.upload_tiles();
  }
  F.R.out_limit = c_cap;
  int32_t counts[kNumBins];
  if (reuse_bins) {
    std::copy(sym_counts, sym_counts + kNumBins, counts);
  } else {
    F.bin(s, 0);
    F.read_counts(s, counts);
  }
  for (int b = 1; b < kNumBins; b++) F.run_bin(s, 1, b, counts[b]);
  const unsigned g = div_up(m + 1, 256);
  if (ptr_width == 8)
    hipLaunchKernelGGL(k_rowptr<int64_t>, dim3(g), dim3(256), 0, s, F.R.off, (int64_t)0, m,
                       static_cast<int64_t*>(row_ptr));
  else
    hipLaunchKernelGGL(k_rowptr<int32_t>, dim3(g), dim3(256), 0, s, F.R.off, (int64_t)0, m,
                       static_cast<int32_t*>(row_ptr));
  BCLX_HIP(hipGetLastError());
}

template struct TwoPhase<float, int32_t>;
template struct TwoPhase<float, int64_t>;
template struct TwoPhase<double, int32_t>;
template struct TwoPhase<double, int64_t>;

struct PlanBase {
  virtual ~PlanBase() = default;
  virtual int64_t symbolic() = 0;
  virtual void numeric(bclx_csr* C) = 0;
  virtual void rows(int64_t* prod, int32_t* nnz) = 0;
};

template <typename T, typename RP>
struct TwoPhasePlan : PlanBase {
  TwoPhase<T, RP> tp;
  TwoPhasePlan(bclx_handle* h, const bclx_csr* A, const bclx_csr* B) : tp(h, make_tile<T, RP>(A, B), BinTuning{}) {}
  int64_t symbolic() override { return tp.symbolic(); }
  void numeric(bclx_csr* C) override {
    BCLX_REQUIRE(C && C->row_ptr, BCLX_ERR_INVALID, "C or C->row_ptr is null");
    BCLX_REQUIRE(C->m == tp.m, BCLX_ERR_INVALID, "C has the wrong number of rows");
    BCLX_REQUIRE(C->ptr_width == 4 || C->ptr_width == 8, BCLX_ERR_INVALID, "C->ptr_width must be 4 or 8");
    BCLX_REQUIRE(C->dtype == (std::is_same<T, float>::value ? BCLX_F32 : BCLX_F64), BCLX_ERR_INVALID,
                 "C dtype must match A and B");
    // a result that cannot be indexed by the requested row_ptr is refused before its arrays are looked at
    BCLX_REQUIRE(tp.nnz_c < 0 || C->ptr_width == 8 || tp.nnz_c <= INT32_MAX, BCLX_ERR_OVERFLOW,
                 "nnz(C) = " + std::to_string(tp.nnz_c) + " does not fit an int32 row_ptr; request ptr_width 8");
    BCLX_REQUIRE(tp.nnz_c == 0 || (C->col_idx && C->vals), BCLX_ERR_INVALID, "C->col_idx / C->vals are null");
    tp.numeric(C->col_idx, static_cast<T*>(C->vals), C->nnz, C->row_ptr, C->ptr_width);
    C->nnz = tp.nnz_c;
  }
  void rows(int64_t* prod, int32_t* nnz) override {
    if (prod) BCLX_HIP(hipMemcpy(prod, tp.F.R.prod, sizeof(int64_t) * tp.m, hipMemcpyDeviceToHost));
    if (nnz) BCLX_HIP(hipMemcpy(nnz, tp.F.R.nnz, sizeof(int32_t) * tp.m, hipMemcpyDeviceToHost));
  }
};

}  // namespace bclx

struct bclx_spgemm_plan {
  bclx_handle* h;
  std::unique_ptr<bclx::PlanBase> impl;
};

using namespace bclx;

extern "C" {

bclx_status bclx_spgemm_symbolic(bclx_handle* h, const bclx_csr* A, const bclx_csr* B, bclx_spgemm_plan** plan,
                                 int64_t* nnz_c) {
  return guarded([&] {
    BCLX_REQUIRE(h && plan && nnz_c, BCLX_ERR_INVALID, "null handle / plan / nnz_c");
    check_pair(A, B);
    set_device(h);
    auto p = std::make_unique<bclx_spgemm_plan>();
    p->h = h;
    const bool f64 = A->dtype == BCLX_F64;
    const bool w8 = A->ptr_width == 8;
    if (!f64 && !w8) p->impl = std::make_unique<TwoPhasePlan<float, int32_t>>(h, A, B);
    else if (!f64 && w8) p->impl = std::make_unique<TwoPhasePlan<float, int64_t>>(h, A, B);
    else if (f64 && !w8) p->impl = std::make_unique<TwoPhasePlan<double, int32_t>>(h, A, B);
    else p->impl = std::make_unique<TwoPhasePlan<double, int64_t>>(h, A, B);
    *nnz_c = p->impl->symbolic();
    *plan = p.release();
  });
}

bclx_status bclx_spgemm(bclx_handle* h, const bclx_csr* A, const bclx_csr* B, bclx_csr* C) {
  return guarded([&] {
    BCLX_REQUIRE(h && C, BCLX_ERR_INVALID, "null handle or C");
    BCLX_REQUIRE(C->ptr_width == 4 || C->ptr_width == 8, BCLX_ERR_INVALID, "C->ptr_width must be 4 or 8");
    check_pair(A, B);
    set_device(h);
    std::unique_ptr<PlanBase> plan;
    const bool f64 = A->dtype == BCLX_F64;
    const bool w8 = A->ptr_width == 8;
    if (!f64 && !w8) plan = std::make_unique<TwoPhasePlan<float, int32_t>>(h, A, B);
    else if (!f64 && w8) plan = std::make_unique<TwoPhasePlan<float, int64_t>>(h, A, B);
    else if (f64 && !w8) plan = std::make_unique<TwoPhasePlan<double, int32_t>>(h, A, B);
    else plan = std::make_unique<TwoPhasePlan<double, int64_t>>(h, A, B);
    const int64_t nnz = plan->symbolic();
    BCLX_REQUIRE(C->ptr_width == 8 || nnz <= INT32_MAX, BCLX_ERR_OVERFLOW,
                 "nnz(C) = " + std::to_string(nnz) + " does not fit an int32 row_ptr; request ptr_width 8");
    // the outputs belong to the caller: they are released here only if this call fails
    const size_t vsz = f64 ? 8 : 4;
    DevBuf rp, col, val;
    rp.alloc((size_t)C->ptr_width * (size_t)(A->m + 1));
    col.alloc(4 * (size_t)std::max<int64_t>(nnz, 1));
    val.alloc(vsz * (size_t)std::max<int64_t>(nnz, 1));
    bclx_csr out{A->m, B->n, nnz, C->ptr_width, A->dtype, rp.p, col.as<int32_t>(), val.p};
    plan->numeric(&out);
    BCLX_HIP(hipStreamSynchronize(h->stream));
    *C = out;
    rp.p = col.p = val.p = nullptr;  // handed over
  });
}

bclx_status bclx_spgemm_numeric(bclx_spgemm_plan* plan, bclx_csr* C) {
  return guarded([&] {
    BCLX_REQUIRE(plan, BCLX_ERR_INVALID, "null plan");
    set_device(plan->h);
    plan->impl->numeric(C);
  });
}

bclx_status bclx_spgemm_plan_rows(bclx_spgemm_plan* plan, int64_t* h_products, int32_t* h_row_nnz) {
  return guarded([&] {
    BCLX_REQUIRE(plan, BCLX_ERR_INVALID, "null plan");
    set_device(plan->h);
    BCLX_HIP(hipStreamSynchronize(plan->h->stream));
    plan->impl->rows(h_products, h_row_nnz);
  });
}

bclx_status bclx_spgemm_plan_destroy(bclx_spgemm_plan* plan) {
  return guarded([&] {
    if (!plan) return;
    set_device(plan->h);
    (void)hipStreamSynchronize(plan->h->stream);
    delete plan;
  });
}

}  // extern "C"
