// SpGEMM host-side building blocks shared by the two-phase path and the batched stage pipeline.
#pragma once

#include <memory>
#include <vector>

#include "internal.hpp"
#include "spgemm_kernels.hpp"

namespace bclx {

void fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v);
void fill_u64(hipStream_t s, unsigned long long* p, int64_t n, unsigned long long v);
int rows_per_block(int bin);
// off[0..n] = exclusive scan of cnt[0..n) (off[n] = total, also written to *total when non-null);
// partial needs scan_partials(n) entries.
int64_t scan_partials(int64_t n);
void scan_counts(hipStream_t s, const int32_t* cnt, int64_t n, int64_t* off, long long* partial, int64_t* total);
void check_pair(const bclx_csr* A, const bclx_csr* B);

template <typename T, typename RP>
TileDev<T, RP> make_tile(const bclx_csr* A, const bclx_csr* B);

// Device workspace over the flat-row space of one batch of products.
template <typename T, typename RP>
struct FlatRows {
  bclx_handle* h = nullptr;
  std::vector<TileDev<T, RP>> tiles;
  int ntiles = 0;
  int64_t nflat = 0, nblocks = 0, max_n = 1;
  Arena arena;
  RowsDev<T, RP> R{};
  long long* scan_partial = nullptr;
  int64_t* scan_total = nullptr;
  DevBuf gbits;  // per-workgroup column bitmaps of the wide-tile heavy kernel (kept zeroed between rows)

  void setup(bclx_handle* h, const std::vector<TileDev<T, RP>>& tiles, bool with_cap);
  void upload_tiles();
  void reset_counters(hipStream_t s);
  // pred_min >= 0 (learned path): also write cap = min(prod, n) and list the rows with more than pred_min products
  // with_stats: also count products (stats[0]) and rows bound for the heavy bin (stats[5])
  void analyze(hipStream_t s, bool do_bin, long long pred_min = -1, bool with_stats = true);
  void bin(hipStream_t s, int src);
  // exclusive scan of cnt over all flat rows into off (nflat + 1 entries), plus off_add; total -> *scan_total
  void scan(hipStream_t s, const int32_t* cnt, int64_t* off, int64_t off_add);
  void run_bin(hipStream_t s, int mode, int b, int64_t count_hint);
  // max_out >= dst_off[nflat]; chunk_row holds div_up(max_out, kCompactChunk) entries
  void compact(hipStream_t s, const int64_t* dst_off, int* dst_col, T* dst_val, const int* src_col, const T* src_val,
               int64_t max_out, int32_t* chunk_row);
  // heavy rows: allocate the zeroed accumulators of the heavy kernel (before any graph capture)
  void prepare_heavy();
  bool heavy_lds_bitmap() const;  // wide-tile heavy kernel: bitmap in LDS (else per-workgroup global bitmaps)
  int64_t read_stat(int q);
  void read_counts(hipStream_t s, int32_t* dst);  // kNumBins bin counts (synchronises s)
};

// The two-phase product C = A * B of one tile pair (symbolic, then numeric into caller-allocated C).
template <typename T, typename RP>
struct TwoPhase {
  bclx_handle* h;
  FlatRows<T, RP> F;
  int64_t nnz_c = -1;
  int64_t m = 0;
  // bins of the symbolic pass; when no row hashed, the numeric pass reuses them (sorted rows are binned by their
  // products alone)
  int32_t sym_counts[kNumBins] = {0};
  bool reuse_bins = false;
  // scratch / scratch_bytes: device memory the workspace is laid out in instead of being allocated
  TwoPhase(bclx_handle* h_, const TileDev<T, RP>& tile, const BinTuning& tn, void* scratch = nullptr,
           size_t scratch_bytes = 0)
      : h(h_) {
    std::vector<TileDev<T, RP>> tiles{tile};
    if (scratch) F.arena.use_external(scratch, scratch_bytes);
    F.setup(h, tiles, false);
    F.R.tn = tn;
    m = tile.m;
  }
  int64_t symbolic();
  // writes col/val into c_col/c_val (capacity c_cap entries) and row_ptr (m+1 entries of 4 or 8 bytes)
  void numeric(int* c_col, T* c_val, int64_t c_cap, void* row_ptr, int ptr_width);
};


}  // namespace bclx
