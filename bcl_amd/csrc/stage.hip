// Batched SUMMA stage: the npairs local products C_p = A_p * B_p of one stage of BCL::gemm(SPMatrix, SPMatrix,
// SPMatrix) (bcl/containers/algorithms/spgemm.hpp:181-242, the `my_a.dot(my_b)` of every owner for one k).
//
//   mode 0  plain: every product runs the two-phase path with its own launches and host synchronisation, the way
//           a per-tile library call (spgemm_cusparse, cusparse_util.cuh:269-514) is used by the reference driver.
//   mode 2  batched exact: one pipeline over the flat rows of all products: analysis, symbolic, scan, numeric.
//   mode 1  batched learned: analysis, MLP-predicted row capacities in place of the symbolic pass, numeric into
//           predicted slots, exact one-pass redo of the rows that overflowed their slot, compaction to CSR.
// Modes 1 and 2 are sized once at plan time by an eager run and then captured in a HIP graph (LDS bins run as
// parallel graph branches), so bclx_stage_execute is one graph launch.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "mlp.hpp"
#include "spgemm.hpp"

namespace bclx {

constexpr int kPredictMinProducts = 32;  // rows with at most this many products keep the exact bound min(prod, n)
constexpr int kPredictBlock = 256;

// Predicted capacities for the rows the analysis pass listed (more than kPredictMinProducts products); every other
// row keeps the exact bound the analysis wrote. The list is dense, so every lane of a wavefront runs the network.
template <typename T, typename RP>
__global__ __launch_bounds__(kPredictBlock) void k_predict(RowsDev<T, RP> R, const float* params, float margin,
                                                           int slack) {
  __shared__ float P[kNParam];
  for (int q = threadIdx.x; q < kNParam; q += blockDim.x) P[q] = params[q];
  __syncthreads();
  const int cnt = *R.pred_count;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&R.stats[4], (unsigned long long)cnt);
  for (int idx = blockIdx.x * kPredictBlock + threadIdx.x; idx < cnt; idx += gridDim.x * kPredictBlock) {
    const int fr = R.pred_items[idx];
    const TileDev<T, RP> tile = R.tiles[R.blk_tile[fr >> 8]];
    const int64_t i = fr - tile.flat0;
    const long long prod = R.prod[fr];
    const long long ub = prod < (long long)tile.n ? prod : (long long)tile.n;
    float f[kFeat];
    row_features<RP>(tile.a_rp, tile.a_col, tile.b_rp, tile.b_col, i, tile.n, 0, tile.avg_a, tile.avg_b, prod, f);
    const float y = mlp_forward_row(P, f);
    const double pred = exp2((double)fminf(fmaxf(y, 0.f), 40.f)) - 1.0;
    if (idx < R.pred_val_cap) R.pred_val[idx] = (float)pred;
    const long long c = (long long)ceil(pred * (1.0 + (double)margin)) + slack;
    long long cap = c < ub ? c : ub;
    if (cap < 1) cap = 1;
    R.cap[fr] = (int32_t)cap;
  }
}

// Predicted rows -> bins by their predicted capacity (every other row was binned by the analysis pass).
template <typename T, typename RP>
__global__ __launch_bounds__(256) void k_bin_predicted(RowsDev<T, RP> R) {
  // per block: count in LDS, one global atomic per bin, then place (the rows of a block share few bins)
  __shared__ int s_cnt[kNumBins], s_base[kNumBins];
  const int cnt = *R.pred_count;
  for (int i0 = blockIdx.x * blockDim.x; i0 < cnt; i0 += gridDim.x * blockDim.x) {
    if (threadIdx.x < kNumBins) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int idx = i0 + threadIdx.x;
    int fr = 0, b = 0, pos = 0;
    if (idx < cnt) {
      fr = R.pred_items[idx];
      const TileDev<T, RP> tile = R.tiles[R.blk_tile[fr >> 8]];
      b = bin_for(R.cap[fr], R.prod[fr], R.tn, tile.n);
      pos = atomicAdd(&s_cnt[b], 1);
    }
    __syncthreads();
    if (threadIdx.x < kNumBins && s_cnt[threadIdx.x] > 0)
      s_base[threadIdx.x] = atomicAdd(R.count_of(threadIdx.x), s_cnt[threadIdx.x]);
    __syncthreads();
    if (idx < cnt) R.bin_items[(int64_t)b * R.bin_stride + s_base[b] + pos] = fr;
    __syncthreads();
  }
}

// Overflow rows of the predicted pass -> LDS bins sized by the exact upper bound min(prod, n).
template <typename T, typename RP>
__global__ __launch_bounds__(256) void k_ovf_bin(RowsDev<T, RP> R) {
  const int cnt = *R.ovf_count;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < cnt; idx += gridDim.x * blockDim.x) {
    const int fr = R.ovf_items[idx];
    const TileDev<T, RP> tile = R.tiles[R.blk_tile[fr >> 8]];
    const long long prod = R.prod[fr];
    const long long ub = prod < (long long)tile.n ? prod : (long long)tile.n;
    const int b = bin_for(ub, prod, R.tn, tile.n);
    const int pos = atomicAdd(R.count_of(b), 1);
    R.bin_items[(int64_t)b * R.bin_stride + pos] = fr;
  }
}

// Flags stats[3] when the predicted slots of this execution need more room than the slot region the plan sized
// (inputs whose structure grew since the plan): rows past the region would overlap the bump-allocated ones.
__global__ void k_check_slots(const int64_t* total, long long region, unsigned long long* flag) {
  if (threadIdx.x == 0 && *total > region) *flag = 1ull;
}

// Flags stats[3] when rows were binned into a bin that this pipeline does not launch (empty when the plan was sized,
// so left out of the graph): inputs whose structure changed since the plan would otherwise lose those rows.
__global__ void k_check_bins(const int32_t* bin_count, unsigned launched, unsigned long long* flag) {
  const int b = threadIdx.x;
  if (b > 0 && b < kNumBins && !((launched >> b) & 1u) && bin_count[b * kCountStride] > 0) *flag = 1ull;
}

// Per-product row_ptr (int64, starting at 0) from flat offsets: product t's row_ptr lives at rp[flat0 + t ...].
template <typename T, typename RP>
__global__ __launch_bounds__(kRowBlock) void k_tile_rowptr(RowsDev<T, RP> R, const int64_t* off, int64_t* rp) {
  const int64_t fr = (int64_t)blockIdx.x * kRowBlock + threadIdx.x;
  const int t = R.blk_tile[blockIdx.x];
  const TileDev<T, RP> tile = R.tiles[t];
  const int64_t i = fr - tile.flat0;
  const int64_t base = off[tile.flat0];
  if (i < tile.m) rp[fr + t] = off[fr] - base;
  if (i == tile.m) rp[fr + t] = off[fr] - base;  // i == m is inside the padding, or the first row of the next block
  if (tile.m % kRowBlock == 0 && i == tile.m - 1) rp[fr + t + 1] = off[fr + 1] - base;
  if (tile.m == 0 && i == 0) rp[fr + t] = 0;
}

// Capacity and exact count of the listed rows: out[2 q] = cap, out[2 q + 1] = nnz of flat row pred_items[q].
template <typename T, typename RP>
__global__ void k_gather_pred(RowsDev<T, RP> R, int cnt, int32_t* out) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
    const int fr = R.pred_items[q];
    out[2 * q] = R.cap[fr];
    out[2 * q + 1] = R.nnz[fr];
  }
}

struct StageBase {
  virtual ~StageBase() = default;
  virtual void execute() = 0;
  virtual void result(int p, bclx_csr* C) = 0;
  virtual void stats(int64_t* out, int n) = 0;
  virtual void predictions(int64_t max, int32_t* product, int32_t* row, float* pred, int32_t* cap, int32_t* nnz,
                           int64_t* count) = 0;
};

template <typename T, typename RP>
struct StagePlan : StageBase {
  bclx_handle* h;
  int mode, npairs;
  bclx_stage_opts opts{};
  BinTuning tn;
  int bin_streams = 1;  // graph branches the bins run on (1: one after another)
  // mode 0
  std::vector<std::unique_ptr<TwoPhase<T, RP>>> per;  // one two-phase workspace per product
  std::vector<DevBuf> c_rp, c_col, c_val;
  std::vector<int64_t> c_nnz;  // capacity of each product's output (its nnz when the plan was made)
  std::vector<int64_t> c_len;  // nnz of each product in the last execution
  // modes 1 / 2
  FlatRows<T, RP> F;
  DevBuf params, arena_col, arena_val, final_col, final_val, final_off, tile_rp, chunk_row, pred_vals;
  int64_t arena_cap = 0, final_cap = 0, main_cap = 0;
  int32_t counts1[kNumBins] = {0}, counts2[kNumBins] = {0};
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> ev;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap_stream = nullptr;
  std::vector<int64_t> host_stats;

  StagePlan(bclx_handle* h_, int np, const bclx_csr* A, const bclx_csr* B, const bclx_stage_opts* o)
      : h(h_), npairs(np) {
    if (o) opts = *o;
    mode = opts.mode;
    if (opts.h_tuning) {
      tn.work_per_lane = std::max(1, opts.h_tuning[0]);
      tn.heavy_need = std::min(std::max(64, opts.h_tuning[1]), 4096);
      tn.load_shift = std::min(std::max(1, opts.h_tuning[2]), 2);
      tn.esc_min = std::max(0, opts.h_tuning[3]);
      tn.esc_max = std::min(std::max(0, opts.h_tuning[4]), kEscMaxProducts);
      bin_streams = std::min(std::max(1, opts.h_tuning[5]), 8);
    }
    std::vector<TileDev<T, RP>> tiles;
    for (int p = 0; p < np; p++) tiles.push_back(make_tile<T, RP>(&A[p], &B[p]));
    if (mode == 0) {
      setup_plain(tiles);
    } else {
      for (int q = 0; q < bin_streams; q++) {
        hipStream_t s;
        BCLX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        side.push_back(s);
      }
      for (int q = 0; q < 2 * (int)side.size() + 2; q++) {
        hipEvent_t e;
        BCLX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ev.push_back(e);
      }
      setup_batched(tiles);
    }
  }

  ~StagePlan() override {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto s : side) (void)hipStreamDestroy(s);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (auto e : ev) (void)hipEventDestroy(e);
  }

  // ---- mode 0 ----------------------------------------------------------------------------------------------
  void plain_product(int p) {
    TwoPhase<T, RP>& W = *per[p];
    const int64_t tot = W.symbolic();
    c_len[p] = 0;
    BCLX_REQUIRE(tot <= c_nnz[p], BCLX_ERR_OVERFLOW, "stage product grew since the plan was made");
    W.numeric(c_col[p].as<int>(), c_val[p].as<T>(), c_nnz[p], c_rp[p].p, 8);
    c_len[p] = tot;
  }

  void setup_plain(const std::vector<TileDev<T, RP>>& tiles) {
    c_rp.resize(npairs);
    c_col.resize(npairs);
    c_val.resize(npairs);
    c_nnz.assign(npairs, 0);
    c_len.assign(npairs, 0);
    for (int p = 0; p < npairs; p++) {
      per.push_back(std::make_unique<TwoPhase<T, RP>>(h, tiles[p], tn));
      const int64_t tot = per.back()->symbolic();  // sizes the outputs
      c_nnz[p] = tot;
      c_len[p] = tot;
      c_col[p].alloc(sizeof(int) * std::max<int64_t>(tot, 1));
      c_val[p].alloc(sizeof(T) * std::max<int64_t>(tot, 1));
      c_rp[p].alloc(sizeof(int64_t) * (tiles[p].m + 1));
    }
    BCLX_HIP(hipStreamSynchronize(h->stream));
  }

  // ---- modes 1 / 2 ------------------------------------------------------------------------------------------
  void fork(hipStream_t s) {
    BCLX_HIP(hipEventRecord(ev[0], s));
    for (auto q : side) BCLX_HIP(hipStreamWaitEvent(q, ev[0], 0));
  }
  void join(hipStream_t s) {
    for (size_t q = 0; q < side.size(); q++) {
      BCLX_HIP(hipEventRecord(ev[1 + q], side[q]));
      BCLX_HIP(hipStreamWaitEvent(s, ev[1 + q], 0));
    }
  }
  // Run the LDS bins and the heavy bin of `mode_k` on the side streams (largest expected first).
  void run_bins(hipStream_t s, int mode_k, const int32_t* counts, bool persistent) {
    unsigned launched = 0;
    for (int b = 1; b < kNumBins; b++)
      if (counts[b] > 0) launched |= 1u << b;
    hipLaunchKernelGGL(k_check_bins, dim3(1), dim3(64), 0, s, (const int32_t*)F.R.bin_count, launched, F.R.stats + 3);
    BCLX_HIP(hipGetLastError());
    fork(s);
    int q = 0;
    if (counts[kHeavyBin] > 0) F.run_bin(side[q++ % side.size()], mode_k, kHeavyBin, persistent ? -1 : counts[kHeavyBin]);
    for (int b = kEscBin0 + kEscBins - 1; b >= kEscBin0; b--) {
      if (counts[b] == 0) continue;
      F.run_bin(side[q++ % side.size()], mode_k, b, counts[b]);
    }
    if (counts[kSmallBin] > 0) F.run_bin(side[q++ % side.size()], mode_k, kSmallBin, counts[kSmallBin]);
    for (int b = 9; b >= 1; b--) {
      if (counts[b] == 0) continue;
      F.run_bin(side[q++ % side.size()], mode_k, b, counts[b]);
    }
    join(s);
  }

  // Analysis with capacities (exact bound everywhere) and binning of every row that is not predicted, then the
  // network on the listed rows (when weights are given) and their binning by the predicted capacity.
  void predict(hipStream_t s, bool with_stats) {
    const bool use_mlp = opts.h_mlp_params != nullptr;
    fill_i32(s, F.R.bin_count, kNumBins * kCountStride, 0);
    fill_i32(s, F.R.pred_count, 1, 0);
    F.analyze(s, true, use_mlp ? (long long)kPredictMinProducts : LLONG_MAX, with_stats);
    if (use_mlp) {
      const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(div_up(F.nflat, kPredictBlock),
                                                                            (int64_t)4 * h->num_cus));
      hipLaunchKernelGGL((k_predict<T, RP>), dim3(nb), dim3(kPredictBlock), 0, s, F.R, params.as<float>(),
                         opts.margin, opts.slack);
      hipLaunchKernelGGL((k_bin_predicted<T, RP>), dim3(std::max<unsigned>(1, h->num_cus)), dim3(256), 0, s, F.R);
      BCLX_HIP(hipGetLastError());
    }
  }

  void enqueue(hipStream_t s, bool persistent) {
    F.reset_counters(s);
    if (mode == 1) {
      fill_i32(s, F.R.nnz, F.nflat, 0);
      fill_u64(s, F.R.cursor, 1, (unsigned long long)main_cap);
      predict(s, false);  // products were counted by the sizing run
      F.scan(s, F.R.cap, F.R.off, 0);
      hipLaunchKernelGGL(k_check_slots, dim3(1), dim3(64), 0, s, (const int64_t*)F.scan_total, (long long)main_cap,
                         F.R.stats + 3);
      BCLX_HIP(hipGetLastError());
      run_bins(s, 2, counts1, persistent);
      fill_i32(s, F.R.bin_count, kNumBins * kCountStride, 0);
      hipLaunchKernelGGL((k_ovf_bin<T, RP>), dim3(std::max<unsigned>(1, std::min<unsigned>(div_up(counts_ovf, 256), 1024))),
                         dim3(256), 0, s, F.R);
      BCLX_HIP(hipGetLastError());
      run_bins(s, 3, counts2, persistent);
      F.scan(s, F.R.nnz, final_off.as<int64_t>(), 0);
      F.compact(s, final_off.as<int64_t>(), final_col.as<int>(), final_val.as<T>(), arena_col.as<int>(),
                arena_val.as<T>(), final_cap, chunk_row.as<int32_t>());
    } else {
      F.analyze(s, true, -1, false);
      run_bins(s, 0, counts1, persistent);
      F.scan(s, F.R.nnz, F.R.off, 0);
      F.bin(s, 0);
      run_bins(s, 1, counts2, persistent);
    }
    const int64_t* off = mode == 1 ? final_off.as<int64_t>() : F.R.off;
    hipLaunchKernelGGL((k_tile_rowptr<T, RP>), dim3((unsigned)F.nblocks), dim3(kRowBlock), 0, s, F.R, off,
                       tile_rp.as<int64_t>());
    BCLX_HIP(hipGetLastError());
  }

  unsigned counts_ovf = 0;
  bool has_heavy = false;
  int64_t products = 0;

  // after the sizing run's analysis: window indexes and scratch for the rows that can take the heavy path
  void setup_heavy(hipStream_t s) {
    (void)s;
    products = F.read_stat(0);  // the sizing run's analysis counts products; graph replays do not
    has_heavy = F.read_stat(5) > 0;
    if (has_heavy) F.prepare_heavy();
  }

  void read_counts(int32_t* dst) {
    F.read_counts(h->stream, dst);
  }

  int64_t read_stat(int q) {
    unsigned long long v = 0;
    BCLX_HIP(hipMemcpyAsync(&v, F.R.stats + q, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    BCLX_HIP(hipStreamSynchronize(h->stream));
    return (int64_t)v;
  }

  // Eager sizing run: the same steps as enqueue, with host reads of the bin counts and totals in between.
  void sizing_run() {
    hipStream_t s = h->stream;
    F.reset_counters(s);
    if (mode == 1) {
      fill_i32(s, F.R.nnz, F.nflat, 0);
      predict(s, true);
      setup_heavy(s);
      if (opts.h_mlp_params) {  // room for the raw predictions of the listed rows (read by bclx_stage_predictions)
        int32_t npred = 0;
        BCLX_HIP(hipMemcpyAsync(&npred, F.R.pred_count, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        BCLX_HIP(hipStreamSynchronize(s));
        const int64_t room = std::min<int64_t>(F.nflat, (int64_t)npred + npred / 4 + 1024);
        pred_vals.alloc(sizeof(float) * (size_t)room);
        F.R.pred_val = pred_vals.as<float>();
        F.R.pred_val_cap = (int32_t)room;
      }
      F.scan(s, F.R.cap, F.R.off, 0);
      int64_t tot_cap = 0;
      BCLX_HIP(hipMemcpyAsync(&tot_cap, F.scan_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
      BCLX_HIP(hipStreamSynchronize(s));
      main_cap = tot_cap;
      // Room for rows that overflow their predicted slot; grown and re-run when this input needs more.
      int64_t reserve = std::max<int64_t>(tot_cap / 8, 1 << 20);
      bool ok = false;
      for (int attempt = 0; attempt < 4 && !ok; attempt++) {
        arena_cap = main_cap + reserve;
        arena_col.alloc(sizeof(int) * arena_cap);
        arena_val.alloc(sizeof(T) * arena_cap);
        for (auto& t : F.tiles) {
          t.c_col = arena_col.as<int>();
          t.c_val = arena_val.as<T>();
        }
        F.upload_tiles();
        F.R.out_limit = arena_cap;
        if (attempt > 0) {  // the previous attempt's overflow redo re-binned: analyse and bin again (same capacities)
          F.reset_counters(s);
          predict(s, false);
        }
        fill_u64(s, F.R.cursor, 1, (unsigned long long)main_cap);
        fill_u64(s, F.R.stats + 2, 2, 0);
        fill_i32(s, F.R.ovf_count, 1, 0);
        fill_i32(s, F.R.nnz, F.nflat, 0);
        read_counts(counts1);
        for (int b = 1; b < kNumBins; b++) F.run_bin(s, 2, b, counts1[b]);
        int32_t novf = 0;
        BCLX_HIP(hipMemcpyAsync(&novf, F.R.ovf_count, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        BCLX_HIP(hipStreamSynchronize(s));
        counts_ovf = (unsigned)novf;
        fill_i32(s, F.R.bin_count, kNumBins * kCountStride, 0);
        hipLaunchKernelGGL((k_ovf_bin<T, RP>), dim3(std::max<unsigned>(1, std::min<unsigned>(div_up(novf, 256), 1024))),
                           dim3(256), 0, s, F.R);
        BCLX_HIP(hipGetLastError());
        read_counts(counts2);
        for (int b = 1; b < kNumBins; b++) F.run_bin(s, 3, b, counts2[b]);
        unsigned long long cur = 0;
        BCLX_HIP(hipMemcpyAsync(&cur, F.R.cursor, sizeof(cur), hipMemcpyDeviceToHost, s));
        BCLX_HIP(hipStreamSynchronize(s));
        // the cursor counts every request, granted or refused, so cur - main_cap is what this input needs
        ok = F.read_stat(3) == 0;
        if (!ok) reserve = ((int64_t)cur - main_cap) + ((int64_t)cur - main_cap) / 16 + 4096;
      }
      BCLX_REQUIRE(ok, BCLX_ERR_OVERFLOW, "could not size the overflow region of the stage");
      F.scan(s, F.R.nnz, F.R.off + 0, 0);  // total nnz only (off is rebuilt by the pipeline)
      int64_t tot = 0;
      BCLX_HIP(hipMemcpyAsync(&tot, F.scan_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
      BCLX_HIP(hipStreamSynchronize(s));
      final_cap = tot + tot / 64 + 4096;
      final_col.alloc(sizeof(int) * final_cap);
      final_val.alloc(sizeof(T) * final_cap);
      chunk_row.alloc(sizeof(int32_t) * (div_up(final_cap, kCompactChunk) + 1));
    } else {
      F.analyze(s, true);
      setup_heavy(s);
      read_counts(counts1);
      for (int b = 1; b < kNumBins; b++) F.run_bin(s, 0, b, counts1[b]);
      F.scan(s, F.R.nnz, F.R.off, 0);
      int64_t tot = 0;
      BCLX_HIP(hipMemcpyAsync(&tot, F.scan_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
      BCLX_HIP(hipStreamSynchronize(s));
      final_cap = tot + tot / 64 + 4096;
      final_col.alloc(sizeof(int) * final_cap);
      final_val.alloc(sizeof(T) * final_cap);
      for (auto& t : F.tiles) {
        t.c_col = final_col.as<int>();
        t.c_val = final_val.as<T>();
      }
      F.upload_tiles();
      F.R.out_limit = final_cap;
      F.bin(s, 0);
      read_counts(counts2);
    }
    BCLX_HIP(hipStreamSynchronize(s));
  }

  void setup_batched(const std::vector<TileDev<T, RP>>& tiles) {
    F.setup(h, tiles, mode == 1);
    F.R.tn = tn;
    params.alloc(sizeof(float) * kNParam);
    if (mode == 1 && opts.h_mlp_params)
      BCLX_HIP(hipMemcpy(params.p, opts.h_mlp_params, sizeof(float) * kNParam, hipMemcpyHostToDevice));
    final_off.alloc(sizeof(int64_t) * (F.nflat + 1));
    tile_rp.alloc(sizeof(int64_t) * (F.nflat + npairs + 1));
    // heavy-row accumulators of tiles wider than the LDS bitmap are allocated before capture
    F.prepare_heavy();
    sizing_run();
    if (opts.use_graph) {
      // capture on a plan-owned stream (the handle's stream may be the legacy default stream, which cannot be
      // captured); the graph is launched on the handle's stream
      BCLX_HIP(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
      hipStream_t s = cap_stream;
      BCLX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(s, false);
      } catch (...) {
        hipGraph_t g;
        (void)hipStreamEndCapture(s, &g);
        throw;
      }
      BCLX_HIP(hipStreamEndCapture(s, &graph));
      BCLX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
  }

  void execute() override {
    if (mode == 0) {
      for (int p = 0; p < npairs; p++) plain_product(p);
      return;
    }
    if (exec) BCLX_HIP(hipGraphLaunch(exec, h->stream));
    else enqueue(h->stream, false);
  }

  void result(int p, bclx_csr* C) override {
    BCLX_REQUIRE(p >= 0 && p < npairs, BCLX_ERR_INVALID, "product index out of range");
    BCLX_HIP(hipStreamSynchronize(h->stream));
    C->ptr_width = 8;
    C->dtype = std::is_same<T, float>::value ? BCLX_F32 : BCLX_F64;
    if (mode == 0) {
      C->m = per[p]->F.tiles[0].m;
      C->n = per[p]->F.tiles[0].n;
      C->nnz = c_len[p];
      C->row_ptr = c_rp[p].p;
      C->col_idx = c_col[p].as<int>();
      C->vals = c_val[p].p;
      return;
    }
    // A row refused for lack of room (inputs whose structure outgrew what the plan was sized for) was not written:
    // no product of this execution is handed out.
    unsigned long long refused = 0;
    BCLX_HIP(hipMemcpy(&refused, F.R.stats + 3, sizeof(refused), hipMemcpyDeviceToHost));
    BCLX_REQUIRE(refused == 0, BCLX_ERR_OVERFLOW,
                 "stage output exceeded the capacity sized at plan time (the inputs' sparsity structure grew since "
                 "the plan was made: make a new plan)");
    const TileDev<T, RP>& t = F.tiles[p];
    const int64_t* off = mode == 1 ? final_off.as<int64_t>() : F.R.off;
    int64_t lo = 0, hi = 0;
    BCLX_HIP(hipMemcpy(&lo, off + t.flat0, sizeof(int64_t), hipMemcpyDeviceToHost));
    BCLX_HIP(hipMemcpy(&hi, off + t.flat0 + t.m, sizeof(int64_t), hipMemcpyDeviceToHost));
    C->m = t.m;
    C->n = t.n;
    C->nnz = hi - lo;
    C->row_ptr = tile_rp.as<int64_t>() + t.flat0 + p;
    C->col_idx = final_col.as<int>() + lo;
    C->vals = final_val.as<T>() + lo;
  }

  void predictions(int64_t max, int32_t* product, int32_t* row, float* pred, int32_t* cap, int32_t* nnz,
                   int64_t* count) override {
    BCLX_REQUIRE(mode == 1 && opts.h_mlp_params, BCLX_ERR_INVALID,
                 "predictions exist only in a learned plan (mode 1) made with network weights");
    BCLX_HIP(hipStreamSynchronize(h->stream));
    int32_t cnt = 0;
    BCLX_HIP(hipMemcpy(&cnt, F.R.pred_count, sizeof(int32_t), hipMemcpyDeviceToHost));
    cnt = std::min<int32_t>(cnt, F.R.pred_val_cap);
    *count = cnt;
    const int64_t k = std::min<int64_t>(cnt, max);
    if (k <= 0) return;
    std::vector<int32_t> items((size_t)k), capnnz(2 * (size_t)k);
    DevBuf tmp(sizeof(int32_t) * 2 * (size_t)cnt);
    hipLaunchKernelGGL((k_gather_pred<T, RP>), dim3(std::max<unsigned>(1, std::min<unsigned>(div_up(cnt, 256), 4096))),
                       dim3(256), 0, h->stream, F.R, cnt, tmp.as<int32_t>());
    BCLX_HIP(hipGetLastError());
    BCLX_HIP(hipMemcpyAsync(items.data(), F.R.pred_items, sizeof(int32_t) * (size_t)k, hipMemcpyDeviceToHost,
                            h->stream));
    BCLX_HIP(hipMemcpyAsync(capnnz.data(), tmp.p, sizeof(int32_t) * 2 * (size_t)k, hipMemcpyDeviceToHost, h->stream));
    if (pred)
      BCLX_HIP(hipMemcpyAsync(pred, F.R.pred_val, sizeof(float) * (size_t)k, hipMemcpyDeviceToHost, h->stream));
    BCLX_HIP(hipStreamSynchronize(h->stream));
    for (int64_t q = 0; q < k; q++) {
      const int64_t fr = items[(size_t)q];
      int t = 0;  // the product holding flat row fr: the last with flat0 <= fr
      int lo = 0, hi = npairs - 1;
      while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        if (F.tiles[mid].flat0 <= fr) { t = mid; lo = mid + 1; }
        else hi = mid - 1;
      }
      if (product) product[q] = t;
      if (row) row[q] = (int32_t)(fr - F.tiles[t].flat0);
      if (cap) cap[q] = capnnz[2 * (size_t)q];
      if (nnz) nnz[q] = capnnz[2 * (size_t)q + 1];
    }
  }

  void stats(int64_t* out, int n) override {
    BCLX_HIP(hipStreamSynchronize(h->stream));
    int64_t st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t rows = 0;
    if (mode == 0) {
      for (int p = 0; p < npairs; p++) {
        unsigned long long v[2];
        BCLX_HIP(hipMemcpy(v, per[p]->F.R.stats, sizeof(v), hipMemcpyDeviceToHost));
        st[0] += (int64_t)v[0];
        st[6] += (int64_t)v[1];
        st[1] += c_len[p];
        st[4] += c_nnz[p];
        rows += per[p]->F.tiles[0].m;
      }
    } else {
      unsigned long long v[5];
      BCLX_HIP(hipMemcpy(v, F.R.stats, sizeof(v), hipMemcpyDeviceToHost));
      const int64_t* off = mode == 1 ? final_off.as<int64_t>() : F.R.off;
      int64_t tot = 0;
      BCLX_HIP(hipMemcpy(&tot, off + F.nflat, sizeof(int64_t), hipMemcpyDeviceToHost));
      st[0] = products;
      st[1] = tot;
      st[2] = (int64_t)v[2];
      st[3] = (int64_t)v[4];
      st[4] = mode == 1 ? main_cap : final_cap;  // learned: sum of the row capacities (the slot region)
      st[6] = (int64_t)v[1];
      BCLX_REQUIRE(v[3] == 0, BCLX_ERR_OVERFLOW,
                   "stage output exceeded the capacity sized at plan time (inputs changed since the plan?)");
      for (auto& t : F.tiles) rows += t.m;
      for (int b = kEscBin0; b < kEscBin0 + kEscBins; b++) st[7] += counts1[b];
      st[7] += counts1[kSmallBin];
    }
    st[5] = rows;
    for (int q = 0; q < n && q < 8; q++) out[q] = st[q];
  }
};

}  // namespace bclx

struct bclx_stage_plan {
  bclx_handle* h;
  std::unique_ptr<bclx::StageBase> impl;
};

using namespace bclx;

extern "C" {

bclx_status bclx_stage_plan_create(bclx_handle* h, int npairs, const bclx_csr* A, const bclx_csr* B,
                                   const bclx_stage_opts* opts, bclx_stage_plan** plan) {
  return guarded([&] {
    BCLX_REQUIRE(h && A && B && plan && npairs > 0, BCLX_ERR_INVALID, "null argument or npairs <= 0");
    const int mode = opts ? opts->mode : 0;
    BCLX_REQUIRE(mode >= 0 && mode <= 2, BCLX_ERR_INVALID, "mode must be 0, 1 or 2");
    for (int p = 0; p < npairs; p++) {
      check_pair(&A[p], &B[p]);
      BCLX_REQUIRE(A[p].dtype == A[0].dtype && A[p].ptr_width == A[0].ptr_width, BCLX_ERR_INVALID,
                   "all products of a stage must share dtype and ptr_width");
    }
    set_device(h);
    auto p = std::make_unique<bclx_stage_plan>();
    p->h = h;
    const bool f64 = A[0].dtype == BCLX_F64, w8 = A[0].ptr_width == 8;
    if (!f64 && !w8) p->impl = std::make_unique<StagePlan<float, int32_t>>(h, npairs, A, B, opts);
    else if (!f64 && w8) p->impl = std::make_unique<StagePlan<float, int64_t>>(h, npairs, A, B, opts);
    else if (f64 && !w8) p->impl = std::make_unique<StagePlan<double, int32_t>>(h, npairs, A, B, opts);
    else p->impl = std::make_unique<StagePlan<double, int64_t>>(h, npairs, A, B, opts);
    *plan = p.release();
  });
}

bclx_status bclx_stage_execute(bclx_stage_plan* plan) {
  return guarded([&] {
    BCLX_REQUIRE(plan, BCLX_ERR_INVALID, "null plan");
    set_device(plan->h);
    plan->impl->execute();
  });
}

bclx_status bclx_stage_result(bclx_stage_plan* plan, int p, bclx_csr* C) {
  return guarded([&] {
    BCLX_REQUIRE(plan && C, BCLX_ERR_INVALID, "null argument");
    set_device(plan->h);
    plan->impl->result(p, C);
  });
}

bclx_status bclx_stage_stats(bclx_stage_plan* plan, int64_t* h_stats, int nstats) {
  return guarded([&] {
    BCLX_REQUIRE(plan && h_stats, BCLX_ERR_INVALID, "null argument");
    set_device(plan->h);
    plan->impl->stats(h_stats, nstats);
  });
}

bclx_status bclx_stage_predictions(bclx_stage_plan* plan, int64_t max, int32_t* h_product, int32_t* h_row,
                                   float* h_pred, int32_t* h_cap, int32_t* h_nnz, int64_t* count) {
  return guarded([&] {
    BCLX_REQUIRE(plan && count && max >= 0, BCLX_ERR_INVALID, "bad argument");
    set_device(plan->h);
    plan->impl->predictions(max, h_product, h_row, h_pred, h_cap, h_nnz, count);
  });
}

bclx_status bclx_stage_plan_destroy(bclx_stage_plan* plan) {
  return guarded([&] {
    if (!plan) return;
    set_device(plan->h);
    (void)hipStreamSynchronize(plan->h->stream);
    delete plan;
  });
}

}  // extern "C"
