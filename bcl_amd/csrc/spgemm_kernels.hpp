// SpGEMM device kernels: row analysis, binning, LDS-hash accumulators (symbolic / numeric / predicted / one-pass),
// heavy-row global accumulators, scans and compaction. Shared by the per-product two-phase path and the batched
// stage pipeline: every kernel works on a "flat row" space that concatenates the rows of all products of a batch,
// each product padded to a multiple of 256 rows, so one 256-row block never straddles two products.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <type_traits>

namespace bclx {

constexpr int kEmptyKey = 0x7fffffff;  // larger than any column index
// 0: empty row; 1..9: LDS hash tables; 10: heavy (global accumulator); 11..17: sort-based (ESC) rows of up to
// 64 * 2^(b - 11) products (bin 11: sorted in registers by 16 lanes per row; 12..14: one wavefront per row; 15..17:
// 2, 4 or 8 wavefronts per row); 18: rows of at most 32 products, sorted in registers by eight lanes per row
constexpr int kNumBins = 19;
constexpr int kHeavyBin = 10;
constexpr int kCountStride = 32;  // bin counters sit 128 bytes apart
constexpr int kEscBin0 = 11;
constexpr int kEscBins = 7;
constexpr int kEscWaveBins = 4;                         // bins 11..14 sort within one wavefront (or less)
constexpr int kEscMaxProducts = 64 << (kEscBins - 1);  // 4096
constexpr int kEscColBits = 19;                         // sort keys pack column (< 2^19) and product index (13 bits)
constexpr int kSmallBin = 18;
constexpr int kSmallMax = 32;       // products per row in the small bin
constexpr int kSmallColBits = 27;   // small-bin keys pack column (< 2^27 - 1) and product index (5 bits)
constexpr int kSmallGroup = 8;      // lanes per small-bin row (four products per lane; 16 and 4 measured slower)
constexpr int kRowBlock = 256;         // flat-row padding and analysis block size
constexpr int kHeavyThreads = 1024;

// Table geometry of LDS bins 1..9: TS = 2^(b+4) slots holds up to TS/2 keys (load <= 1/2); G lanes per row.
// Groups are either part of one wavefront (G <= 64, several rows per block) or the whole block (G >= 256).
__host__ __device__ constexpr int bin_table(int b) { return 1 << (b + 4); }
__host__ __device__ constexpr int bin_group(int b) {
  return b == 1 ? 16 : b == 2 ? 32 : b <= 5 ? 64 : b == 6 ? 256 : b <= 8 ? 512 : 1024;
}

// Tunable binning knobs (chosen by the autotuner; these are the defaults).
struct BinTuning {
  int32_t work_per_lane = 64;     // a row's group must have >= products / work_per_lane lanes
  int32_t heavy_need = 4096;      // rows needing more table entries than this use the global accumulator
  int32_t load_shift = 1;         // table slots >= need << load_shift (1: load <= 1/2)
  int32_t esc_min = 32;           // rows with esc_min < products <= esc_max sort instead of hashing (when the
  int32_t esc_max = 4096;         //   tile is at most 2^19 columns wide); esc_max 0 turns sorting off
};

// ESC bin of a row with `work` products: the smallest 64 * 2^e >= work.
__host__ __device__ inline int esc_bin(int64_t work) {
  int e = 0;
  while ((64LL << e) < work) e++;
  return kEscBin0 + e;
}

// n: the tile's width. Rows of at most 32 products sort in registers (tiles narrower than 2^27 - 1 columns); longer
// rows sort in the ESC bins when the columns fit their packed key (n <= 2^19). esc_max = 0 turns both off.
__host__ __device__ inline int bin_for(int64_t need, int64_t work, const BinTuning& tn, int64_t n) {
  if (work <= 0) return 0;
  if (tn.esc_max > 0 && work <= kSmallMax && n < ((int64_t)1 << kSmallColBits) - 1) return kSmallBin;
  const bool esc_ok = n <= (1 << kEscColBits);
  if (esc_ok && work > tn.esc_min && work <= tn.esc_max && work <= kEscMaxProducts) return esc_bin(work);
  if (need > tn.heavy_need) return kHeavyBin;
  int64_t slots = need << tn.load_shift;
  int bn = 1;
  while (bn < 9 && bin_table(bn) < slots) bn++;
  if (bin_table(bn) < slots) return kHeavyBin;
  int bw = bn;
  while (bw < 9 && (int64_t)bin_group(bw) * tn.work_per_lane < work) bw++;
  return bw;
}

// Home slot of column j in a table of TS = 2^b slots: Fibonacci hashing (top b bits of j * 2^32 / phi), so every
// bit of j moves the slot. (A multiplier-and-mask hash keeps only j's low bits; R-MAT columns have biased bits, so
// their low bits take few distinct values and linear probing would chain.)
template <int TS>
__device__ inline unsigned slot_hash(int j) {
  constexpr int b = __builtin_ctz(TS);
  return ((unsigned)j * 0x9E3779B1u) >> (32 - b);
}

template <typename T, typename RP>
struct TileDev {
  const RP* a_rp;
  const int* a_col;
  const T* a_val;
  const RP* b_rp;
  const int* b_col;
  const T* b_val;
  // output base (positions come from off[]). The kernels write it with nontemporal (streaming) stores, except those
  // of rows of at most 64 products: an entry is written once and not read by this launch, so it need not displace
  // the B rows held in L2.
  int* c_col;
  T* c_val;
  int64_t flat0;  // first flat row of this product
  int32_t m, n;
  float avg_a, avg_b;      // average row lengths of A and B (predictor features)
  int64_t krows;           // rows of B
};

template <typename T, typename RP>
struct RowsDev {
  const TileDev<T, RP>* tiles;
  const uint16_t* blk_tile;  // (flat row >> 8) -> product index
  int64_t* prod;             // products per flat row
  int32_t* nnz;              // nnz(C_i): symbolic count, or measured count
  int32_t* cap;              // predicted capacity (learned path)
  int64_t* off;              // output offset per flat row (nflat + 1 entries)
  int32_t* bin_count;        // kNumBins counters, kCountStride ints apart (one cache line each: no false sharing of
                             // the atomics that fill them)
  __host__ __device__ int32_t* count_of(int b) const { return bin_count + b * kCountStride; }
  int32_t* bin_items;        // kNumBins lists of flat rows, bin_stride apart
  int64_t bin_stride;
  int32_t* ovf_count;        // overflow list (learned path)
  int32_t* ovf_items;
  int32_t* pred_count;       // rows whose capacity the predictor sets (learned path)
  int32_t* long_count;       // rows of more than 32 A entries, listed by the analysis in tmp_items
  int32_t* pred_items;
  float* pred_val;             // the network's raw prediction per listed row (learned path; pred_val_cap entries)
  int32_t pred_val_cap;
  unsigned long long* cursor;  // bump allocator for rows placed outside their predicted slot
  unsigned long long* heavy_queue;  // next heavy row to take (rows are handed out largest first)
  int32_t* tmp_items;               // scratch list for ordering the heavy bin
  unsigned long long* stats;   // [0] products [1] heavy rows [2] overflow rows [3] capacity exceeded [4] predicted
  long long out_limit;         // entries available in the output arrays; writes beyond it are refused
  BinTuning tn;
};

// ------------------------------------------------------------------------------------------------------------
// reductions / scans
// ------------------------------------------------------------------------------------------------------------

template <typename V>
__device__ inline V wave_sum(V v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Inclusive scan within a wave.
template <typename V>
__device__ inline V wave_incl_scan(V v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    V t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Block-wide exclusive scan (BLOCK a multiple of 64, <= 1024). Returns exclusive prefix; *total = block sum.
template <int BLOCK, typename V>
__device__ inline V block_excl_scan(V v, V* total, V* smem /* >= BLOCK/64 + 1 */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int NW = BLOCK / 64;
  V inc = wave_incl_scan(v);
  if (lane == 63) smem[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    V run = 0;
    for (int q = 0; q < NW; q++) {
      V t = smem[q];
      smem[q] = run;
      run += t;
    }
    smem[NW] = run;
  }
  __syncthreads();
  V r = smem[w] + inc - v;
  *total = smem[NW];
  __syncthreads();
  return r;
}

template <int BLOCK, typename V>
__device__ inline V block_sum(V v, V* smem) {
  V t;
  (void)block_excl_scan<BLOCK, V>(v, &t, smem);
  return t;
}

// ------------------------------------------------------------------------------------------------------------
// row analysis: products per row, optional binning by upper bound min(products, n) (symbolic bins)
// ------------------------------------------------------------------------------------------------------------

// 256-row blocks of flat rows per analysis workgroup: each thread analyzes one row of each, with the loads of its
// rows issued together, and the workgroup takes its bin and list positions with one global atomic per bin.
constexpr int kAnaSub = 4;

template <typename T, typename RP>
__global__ __launch_bounds__(kRowBlock) void k_analyze(RowsDev<T, RP> R, int64_t nblocks, int do_bin,
                                                       long long pred_min, int with_stats) {
  __shared__ int s_nlong, s_lbase;
  __shared__ int s_bcount[kNumBins], s_bbase[kNumBins];
  __shared__ int s_npred, s_pbase;
  __shared__ long long s_red[kRowBlock / 64 + 1];
  const int tid = threadIdx.x;
  const int64_t sb0 = (int64_t)blockIdx.x * kAnaSub;
  if (tid == 0) {
    s_nlong = 0;
    s_npred = 0;
  }
  if (tid < kNumBins) s_bcount[tid] = 0;
  __syncthreads();
  const TileDev<T, RP>* tp[kAnaSub];
  bool in[kAnaSub], valid[kAnaSub];
  int64_t i[kAnaSub];
  RP a0[kAnaSub], a1[kAnaSub];
  long long prod[kAnaSub];
  bool lng[kAnaSub];
#pragma unroll
  for (int s = 0; s < kAnaSub; s++) {
    in[s] = sb0 + s < nblocks;
    tp[s] = R.tiles + (in[s] ? R.blk_tile[sb0 + s] : 0);
    i[s] = (sb0 + s) * kRowBlock + tid - tp[s]->flat0;
    valid[s] = in[s] && i[s] >= 0 && i[s] < tp[s]->m;
    a0[s] = valid[s] ? tp[s]->a_rp[i[s]] : 0;
    a1[s] = valid[s] ? tp[s]->a_rp[i[s] + 1] : 0;
    prod[s] = 0;
  }
  // rows of up to 32 entries: four entries of every row at a time, their column loads and then their B-row-pointer
  // loads issued together. Longer rows (R-MAT hubs cluster in a few workgroups) go on a global list that
  // k_analyze_long spreads over the whole GPU, so they do not hold their workgroup.
  bool more = false;
  int lpos[kAnaSub];
#pragma unroll
  for (int s = 0; s < kAnaSub; s++) {
    lng[s] = a1[s] - a0[s] > 32;
    lpos[s] = lng[s] ? atomicAdd(&s_nlong, 1) : 0;
    if (lng[s]) a1[s] = a0[s];
    more = more || a1[s] > a0[s];
  }
  for (RP p0 = 0; more; p0 += 4) {
    int k[kAnaSub][4];
    more = false;
#pragma unroll
    for (int s = 0; s < kAnaSub; s++)
#pragma unroll
      for (int u = 0; u < 4; u++) k[s][u] = a0[s] + p0 + u < a1[s] ? tp[s]->a_col[a0[s] + p0 + u] : -1;
#pragma unroll
    for (int s = 0; s < kAnaSub; s++) {
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (k[s][u] >= 0) prod[s] += (long long)(tp[s]->b_rp[k[s][u] + 1] - tp[s]->b_rp[k[s][u]]);
      more = more || a0[s] + p0 + 4 < a1[s];
    }
  }
  __syncthreads();
  if (tid == 0 && s_nlong > 0) s_lbase = atomicAdd(R.long_count, s_nlong);
  __syncthreads();
  int b0[kAnaSub];
  bool pr[kAnaSub];
  long long ub0[kAnaSub];
#pragma unroll
  for (int s = 0; s < kAnaSub; s++) {
    const int64_t fr = (sb0 + s) * kRowBlock + tid;
    if (lng[s]) {  // finished by k_analyze_long
      R.tmp_items[s_lbase + lpos[s]] = (int32_t)fr;
      valid[s] = false;
      in[s] = false;
    }
    if (in[s]) R.prod[fr] = valid[s] ? prod[s] : 0;
    ub0[s] = prod[s] < (long long)tp[s]->n ? prod[s] : (long long)tp[s]->n;
    b0[s] = valid[s] ? bin_for(ub0[s], prod[s], R.tn, tp[s]->n) : 0;
    if (with_stats && b0[s] == kHeavyBin) atomicAdd(&R.stats[5], 1ull);
    pr[s] = false;
  }
  if (pred_min >= 0) {
    // learned path: every row starts at the exact bound min(prod, n); rows with more than pred_min products that
    // hash (sorted rows are exact in one pass and keep the bound) go on the list whose capacities the predictor
    // then lowers, and are binned after the prediction
    int ppos[kAnaSub];
#pragma unroll
    for (int s = 0; s < kAnaSub; s++) {
      const int64_t fr = (sb0 + s) * kRowBlock + tid;
      if (in[s]) R.cap[fr] = valid[s] ? (int32_t)ub0[s] : 0;
      pr[s] = valid[s] && prod[s] > pred_min && b0[s] < kEscBin0;
      ppos[s] = pr[s] ? atomicAdd(&s_npred, 1) : 0;
    }
    __syncthreads();
    if (tid == 0 && s_npred > 0) s_pbase = atomicAdd(R.pred_count, s_npred);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kAnaSub; s++)
      if (pr[s]) R.pred_items[s_pbase + ppos[s]] = (int32_t)((sb0 + s) * kRowBlock + tid);
  }
  if (do_bin) {
    int b[kAnaSub], pos[kAnaSub];
#pragma unroll
    for (int s = 0; s < kAnaSub; s++) {
      b[s] = pr[s] ? 0 : b0[s];
      pos[s] = b[s] > 0 ? atomicAdd(&s_bcount[b[s]], 1) : 0;
    }
    __syncthreads();
    if (tid > 0 && tid < kNumBins && s_bcount[tid] > 0) s_bbase[tid] = atomicAdd(R.count_of(tid), s_bcount[tid]);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kAnaSub; s++) {
      const int64_t fr = (sb0 + s) * kRowBlock + tid;
      if (b[s] > 0) R.bin_items[b[s] * R.bin_stride + s_bbase[b[s]] + pos[s]] = (int)fr;
      if (in[s] && b[s] == 0 && !pr[s]) R.nnz[fr] = 0;  // empty and padding rows
    }
  }
  if (with_stats) {  // one global atomic per block: only when asked (a plan counts products once)
    long long mine = 0;
#pragma unroll
    for (int s = 0; s < kAnaSub; s++) mine += prod[s];
    long long tot = block_sum<kRowBlock, long long>(mine, s_red);
    if (tid == 0 && tot) atomicAdd(&R.stats[0], (unsigned long long)tot);
  }
}

// The rows k_analyze listed (more than 32 A entries): one wavefront per row, rows strided over the wavefronts (the
// list holds each workgroup's long rows together, so striding spreads neighbouring hubs), eight entries per lane in
// flight, then the same outputs as k_analyze gives a row (products, capacity, predictor list, bin), with per-row
// atomics.
// The loop control is wave-uniform arithmetic and lane 0's writes are one if-block at the end of the body: no other
// lane-0 region in the loop (a queue pop by lane 0 at the top was merged by the compiler with the writes at the
// bottom, and the other lanes then looped on their own on a stale index).
template <typename T, typename RP>
__global__ __launch_bounds__(kRowBlock) void k_analyze_long(RowsDev<T, RP> R, int do_bin, long long pred_min,
                                                            int with_stats) {
  const int lane = threadIdx.x & 63;
  const int nwaves = (int)gridDim.x * (kRowBlock / 64);
  const int cnt = R.long_count[0];
  for (int q = (int)blockIdx.x * (kRowBlock / 64) + (int)(threadIdx.x >> 6); q < cnt; q += nwaves) {
    const int fr = R.tmp_items[q];
    const TileDev<T, RP>* t = R.tiles + R.blk_tile[fr >> 8];
    const int64_t i = fr - t->flat0;
    const RP a0 = t->a_rp[i], a1 = t->a_rp[i + 1];
    long long prod = 0;
    for (RP p = a0 + lane; p < a1; p += 8 * 64) {
      int kk[8];
#pragma unroll
      for (int u = 0; u < 8; u++) kk[u] = p + 64 * u < a1 ? t->a_col[p + 64 * u] : -1;
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (kk[u] >= 0) prod += (long long)(t->b_rp[kk[u] + 1] - t->b_rp[kk[u]]);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) prod += __shfl_xor(prod, d);
    if (lane == 0) {
      R.prod[fr] = prod;
      const long long ub0 = prod < (long long)t->n ? prod : (long long)t->n;
      const int b0 = bin_for(ub0, prod, R.tn, t->n);
      if (with_stats) {
        atomicAdd(&R.stats[0], (unsigned long long)prod);
        if (b0 == kHeavyBin) atomicAdd(&R.stats[5], 1ull);
      }
      bool pr = false;
      if (pred_min >= 0) {
        R.cap[fr] = (int32_t)ub0;
        pr = prod > pred_min && b0 < kEscBin0;
        if (pr) R.pred_items[atomicAdd(R.pred_count, 1)] = fr;
      }
      if (do_bin) {
        const int b = pr ? 0 : b0;
        if (b > 0) R.bin_items[b * R.bin_stride + atomicAdd(R.count_of(b), 1)] = fr;
        if (b == 0 && !pr) R.nnz[fr] = 0;
      }
    }
  }
}

// Bin rows by `need` (src 0: nnz, 1: cap) and products. Padded / empty rows go nowhere.
template <typename T, typename RP>
__global__ __launch_bounds__(kRowBlock) void k_bin(RowsDev<T, RP> R, int src) {
  __shared__ int s_bcount[kNumBins], s_bbase[kNumBins];
  const int tid = threadIdx.x;
  const int64_t fr = (int64_t)blockIdx.x * kRowBlock + tid;
  const TileDev<T, RP> tile = R.tiles[R.blk_tile[blockIdx.x]];
  const int64_t i = fr - tile.flat0;
  const bool valid = i >= 0 && i < tile.m;
  if (tid < kNumBins) s_bcount[tid] = 0;
  __syncthreads();
  int b = 0;
  if (valid) {
    const long long work = R.prod[fr];
    const long long need = src == 0 ? (long long)R.nnz[fr] : (long long)R.cap[fr];
    b = (need == 0 && src == 0) ? 0 : bin_for(need, work, R.tn, tile.n);
  }
  int pos = 0;
  if (b > 0) pos = atomicAdd(&s_bcount[b], 1);
  __syncthreads();
  if (tid > 0 && tid < kNumBins && s_bcount[tid] > 0) s_bbase[tid] = atomicAdd(R.count_of(tid), s_bcount[tid]);
  __syncthreads();
  if (b > 0) R.bin_items[b * R.bin_stride + s_bbase[b] + pos] = (int)fr;
}

// ------------------------------------------------------------------------------------------------------------
// LDS hash accumulators
//   MODE 0 symbolic: count distinct columns -> nnz[fr]
//   MODE 1 numeric: accumulate, sort, write at off[fr] (nnz known)
//   MODE 2 predicted: table sized by cap[fr]; more than cap distinct columns -> overflow list, else write at off[fr]
//   MODE 3 one-pass exact: table sized by the upper bound; count, bump-allocate from *cursor, write
// Values accumulate in float64 (products of fp32 inputs are exact in float64) and round once on output.
// ------------------------------------------------------------------------------------------------------------

// Synchronise the lanes of one row group: a wavefront-level fence when the group lives inside one wavefront, a
// workgroup barrier when the group is the whole block.
template <int G>
__device__ inline void group_sync() {
  if constexpr (G > 64) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
}

// Exclusive scan of one value per lane over a row group; *total = group sum.
template <int G, int BLOCK>
__device__ inline long long group_excl_scan(long long v, long long* total, long long* smem) {
  if constexpr (G > 64) {
    static_assert(G == BLOCK, "multi-wavefront groups must be the whole block");
    return block_excl_scan<BLOCK, long long>(v, total, smem);
  } else {
    const int gl = threadIdx.x & (G - 1);
    long long inc = v;
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
      const long long t = __shfl_up(inc, d, G);
      if (gl >= d) inc += t;
    }
    *total = __shfl(inc, G - 1, G);
    return inc - v;
  }
}

// The product loop of a staged chunk of A entries: entry e's products are B positions e_start[e] + 0 ..
// e_pre[e + 1] - e_pre[e] - 1 (e_pre holds ne + 1 prefix counts, e_pre[ne] = total). Lane l takes products l,
// l + G, ..., U at a time: the U entry lookups (binary searches over the prefixes) advance in lockstep and the U loads
// are issued together, so their latencies overlap. f(j, v) per product (v = a * b in float64; 0 when !NUM).
template <int G, bool NUM, int U, typename T, typename RP, typename PT, typename F>
__device__ inline void enumerate_staged(const TileDev<T, RP>& tile, int ne, long long total, int lane,
                                        const RP* e_start, const PT* e_pre, const T* e_val,
                                        const volatile int* stop, F&& f) {
  int top = 1;  // largest power of two < ne (0 when ne == 1)
  while (top < ne) top <<= 1;
  top >>= 1;
  for (long long t0 = lane; t0 < total; t0 += (long long)U * G) {
    if (stop && *stop) break;
    int lo[U];
#pragma unroll
    for (int u = 0; u < U; u++) lo[u] = 0;
    for (int step = top; step > 0; step >>= 1) {  // entry e with e_pre[e] <= t < e_pre[e + 1]
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int c = lo[u] + step;
        if (c < ne && e_pre[c] <= t0 + (long long)u * G) lo[u] = c;
      }
    }
    int j[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const long long t = t0 + (long long)u * G;
      if (t < total) {
        const RP q = e_start[lo[u]] + (RP)(t - e_pre[lo[u]]);
        j[u] = tile.b_col[q];
        v[u] = NUM ? (double)e_val[lo[u]] * (double)tile.b_val[q] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (t0 + (long long)u * G < total) f(j[u], v[u]);
  }
}

// Load-balanced enumeration of the products of one row: the group stages the B-row ranges of up to G entries of
// A's row in LDS, prefix-sums their lengths and runs enumerate_staged over them. `stop` (optional, LDS) ends the
// enumeration early once set.
template <int G, int BLOCK, bool NUM, typename T, typename RP, typename PT, typename F, int U = 4>
__device__ inline void for_each_product(const TileDev<T, RP>& tile, RP a0, RP a1, int lane, RP* e_start,
                                        PT* e_pre, T* e_val, long long* scan_smem, const volatile int* stop,
                                        F&& f) {
  for (RP c0 = a0; c0 < a1; c0 += G) {
    const int ne = (int)((a1 - c0) < (RP)G ? (a1 - c0) : (RP)G);
    long long len = 0;
    if (lane < ne) {
      const int kk = tile.a_col[c0 + lane];
      const RP st = tile.b_rp[kk];
      len = (long long)(tile.b_rp[kk + 1] - st);
      e_start[lane] = st;
      if (NUM) e_val[lane] = tile.a_val[c0 + lane];
    }
    long long total;
    const long long ex = group_excl_scan<G, BLOCK>(len, &total, scan_smem);
    if (lane < ne) e_pre[lane] = (PT)ex;
    if (lane == 0) e_pre[ne] = (PT)total;
    group_sync<G>();
    enumerate_staged<G, NUM, U>(tile, ne, total, lane, e_start, e_pre, e_val, stop, f);
    group_sync<G>();
  }
}

template <int TS, int G, int BLOCK>
__device__ inline void bitonic_sort_groups(int* keys, double* vals, int lane, bool with_vals) {
  // Sort each group's TS-slot table ascending by key; kEmptyKey sorts last. All threads of the block call this.
  for (int k = 2; k <= TS; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = lane; t < TS / 2; t += G) {
        const int i = 2 * t - (t & (j - 1));
        const int p = i + j;
        const bool asc = (i & k) == 0;
        const int ki = keys[i], kp = keys[p];
        if ((ki > kp) == asc) {
          keys[i] = kp;
          keys[p] = ki;
          if (with_vals) {
            double vi = vals[i];
            vals[i] = vals[p];
            vals[p] = vi;
          }
        }
      }
      if (G > 64) {
        __syncthreads();
      } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      }
    }
  }
}

// Block-wide min / max of an int (BLOCK <= 1024); smem >= BLOCK/64 ints. Contains barriers.
template <int BLOCK>
__device__ inline int block_min(int v, int* smem) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
  if ((threadIdx.x & 63) == 0) smem[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = smem[0];
  for (int q = 1; q < BLOCK / 64; q++) r = min(r, smem[q]);
  __syncthreads();
  return r;
}
template <int BLOCK>
__device__ inline int block_max(int v, int* smem) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
  if ((threadIdx.x & 63) == 0) smem[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = smem[0];
  for (int q = 1; q < BLOCK / 64; q++) r = max(r, smem[q]);
  __syncthreads();
  return r;
}

// LDS for the bitmap ranking of one row: W words of bits over the row's column span plus one prefix count per
// 8 words. kRankWords words cover spans up to 2^19 columns.
constexpr int kRankWords = 16384;
constexpr int kRankBytes = kRankWords * 4 + (kRankWords / 8 + 16) * 4;

template <int TS, int G, int BLOCK, int MODE, typename T, typename RP>
__global__ __launch_bounds__(BLOCK) void k_hash(RowsDev<T, RP> R, int bin) {
  constexpr int NG = BLOCK / G;
  constexpr bool NUM = MODE != 0;
  // Whole-block rows with large tables: the sorted order comes from ranking every key in a bitmap of the row's
  // column span (prefix popcounts) instead of sorting the table; entries go straight to their output position.
  constexpr bool RANK = NUM && NG == 1 && TS >= 2048;
  constexpr int TBYTES = NG * TS * (NUM ? 12 : 4);
  // entry staging of for_each_product: start (RP), prefix (long long), A value (T) per lane, 16-byte aligned
  constexpr int EPRE = ((NG * (G + 1) * 8) + 15) / 16 * 16;
  constexpr int EST = ((NG * G * (int)sizeof(RP)) + 15) / 16 * 16;
  constexpr int EBYTES = EPRE + EST + ((NG * G * (int)sizeof(T)) + 15) / 16 * 16;
  // the staging lives after the table; the rank bitmap later reuses the whole buffer
  constexpr int LBYTES = (RANK && kRankBytes > TBYTES + EBYTES) ? kRankBytes : TBYTES + EBYTES;
  __shared__ __align__(16) unsigned char s_buf[LBYTES];
  __shared__ int s_cnt[NG];
  __shared__ int s_flag[NG];
  __shared__ long long s_base[NG];
  __shared__ int s_red[BLOCK / 64 + 1];
  __shared__ long long s_lscan[BLOCK / 64 + 1];
  long long* s_epre = reinterpret_cast<long long*>(s_buf + TBYTES);
  RP* s_est = reinterpret_cast<RP*>(s_buf + TBYTES + EPRE);
  T* s_eval = reinterpret_cast<T*>(s_buf + TBYTES + EPRE + EST);
  const int g = threadIdx.x / G, lane = threadIdx.x % G;
  int* keys = reinterpret_cast<int*>(s_buf) + g * TS;
  double* vals = reinterpret_cast<double*>(s_buf + NG * TS * 4) + (NUM ? g * TS : 0);
  const int count = *R.count_of(bin);
  const int* items = R.bin_items + (int64_t)bin * R.bin_stride;
  volatile int* vflag = s_flag;

  for (int base = blockIdx.x * NG; base < count; base += gridDim.x * NG) {
    const int idx = base + g;
    const bool active = idx < count;
    const int fr = active ? items[idx] : 0;
    for (int s = lane; s < TS; s += G) {
      keys[s] = kEmptyKey;
      if (NUM) vals[s] = 0.0;
    }
    if (lane == 0) {
      s_cnt[g] = 0;
      s_flag[g] = 0;
    }
    __syncthreads();
    TileDev<T, RP> tile;
    if (active) {
      tile = R.tiles[R.blk_tile[fr >> 8]];
      const int64_t i = fr - tile.flat0;
      const int limit = MODE == 2 ? R.cap[fr] : INT_MAX;
      const RP a0 = tile.a_rp[i], a1 = tile.a_rp[i + 1];
      auto insert = [&](int j, double pv) {
        unsigned h = slot_hash<TS>(j);
        for (int probe = 0; probe < TS; probe++) {
          const int kcur = keys[h];
          if (kcur == j) {
            if (NUM) atomicAdd(&vals[h], pv);
            return;
          }
          if (kcur == kEmptyKey) {
            const int old = atomicCAS(&keys[h], kEmptyKey, j);
            if (old == kEmptyKey) {
              const int c = atomicAdd(&s_cnt[g], 1);
              if (MODE == 2 && c >= limit) vflag[g] = 1;
              if (NUM) atomicAdd(&vals[h], pv);
              return;
            }
            if (old == j) {
              if (NUM) atomicAdd(&vals[h], pv);
              return;
            }
          }
          h = (h + 1) & (TS - 1);
        }
        if (MODE == 2) vflag[g] = 1;  // table full: only possible after overflow
      };
      for_each_product<G, BLOCK, NUM, T, RP>(tile, a0, a1, lane, s_est + g * G, s_epre + g * (G + 1),
                                              s_eval + (NUM ? g * G : 0), s_lscan, MODE == 2 ? vflag + g : nullptr,
                                              insert);
    }
    __syncthreads();
    if (MODE == 0) {
      if (active && lane == 0) R.nnz[fr] = s_cnt[g];
    } else {
      const int cnt = s_cnt[g];
      const bool ovf = MODE == 2 && s_flag[g];
      // A row refused for lack of output room keeps nnz = 0 (and a harmless offset), so later passes that copy
      // nnz entries from its offset (compaction) stay inside the arrays; the refusal is flagged in stats[3].
      if (active && lane == 0) {
        if (MODE == 2) {
          if (ovf) {
            int slot = atomicAdd(R.ovf_count, 1);
            R.ovf_items[slot] = fr;
            atomicAdd(&R.stats[2], 1ull);
          } else {
            const bool fits = R.off[fr] + cnt <= R.out_limit;
            R.nnz[fr] = fits ? cnt : 0;
            if (!fits) R.stats[3] = 1ull;
          }
        } else if (MODE == 3) {
          s_base[g] = (long long)atomicAdd(R.cursor, (unsigned long long)cnt);
          const bool fits = s_base[g] + cnt <= R.out_limit;
          R.off[fr] = fits ? s_base[g] : 0;
          R.nnz[fr] = fits ? cnt : 0;
          if (!fits) R.stats[3] = 1ull;
        }
      }
      bool ranked = false;
      if constexpr (RANK) {
        constexpr int SPT = TS / BLOCK;
        int rk[SPT];
        double rv[SPT];
        int kmin = INT_MAX, kmax = -1;
#pragma unroll
        for (int q = 0; q < SPT; q++) {
          rk[q] = keys[q * BLOCK + threadIdx.x];
          rv[q] = vals[q * BLOCK + threadIdx.x];
          if (rk[q] != kEmptyKey) {
            kmin = min(kmin, rk[q]);
            kmax = max(kmax, rk[q]);
          }
        }
        kmin = block_min<BLOCK>(kmin, s_red);
        kmax = block_max<BLOCK>(kmax, s_red);
        const long long span = kmax >= kmin ? (long long)kmax - kmin + 1 : 1;
        const int W = (int)((span + 31) / 32);
        if (W <= kRankWords && (W * 4 + ((W + 7) / 8 + 1) * 4) <= LBYTES) {
          ranked = true;
          unsigned* bits = reinterpret_cast<unsigned*>(s_buf);
          int* pre = reinterpret_cast<int*>(s_buf + (size_t)W * 4);
          const int NGR = (W + 7) / 8;
          for (int w = threadIdx.x; w < W; w += BLOCK) bits[w] = 0u;
          __syncthreads();
#pragma unroll
          for (int q = 0; q < SPT; q++)
            if (rk[q] != kEmptyKey) {
              const int rel = rk[q] - kmin;
              atomicOr(&bits[rel >> 5], 1u << (rel & 31));
            }
          __syncthreads();
          const int gpt = (NGR + BLOCK - 1) / BLOCK;
          const int g0 = threadIdx.x * gpt, g1 = min(NGR, g0 + gpt);
          int c = 0;
          for (int gg = g0; gg < g1; gg++)
            for (int w = gg * 8; w < min(gg * 8 + 8, W); w++) c += __popc(bits[w]);
          int tot;
          int run = block_excl_scan<BLOCK, int>(c, &tot, s_red);
          for (int gg = g0; gg < g1; gg++) {
            pre[gg] = run;
            for (int w = gg * 8; w < min(gg * 8 + 8, W); w++) run += __popc(bits[w]);
          }
          __syncthreads();
          const long long o = MODE == 3 ? s_base[0] : (active ? R.off[fr] : 0);
          const bool fits = o + cnt <= R.out_limit;
          if (active && !ovf && !fits && threadIdx.x == 0) R.stats[3] = 1ull;
          if (active && !ovf && fits) {
#pragma unroll
            for (int q = 0; q < SPT; q++) {
              if (rk[q] == kEmptyKey) continue;
              const int rel = rk[q] - kmin;
              const int w = rel >> 5;
              int r = pre[w >> 3];
              for (int w2 = (w >> 3) * 8; w2 < w; w2++) r += __popc(bits[w2]);
              r += __popc(bits[w] & ((1u << (rel & 31)) - 1u));
              __builtin_nontemporal_store(rk[q], &tile.c_col[o + r]);
              __builtin_nontemporal_store((T)rv[q], &tile.c_val[o + r]);
            }
          }
        } else {
          __syncthreads();
#pragma unroll
          for (int q = 0; q < SPT; q++) {
            keys[q * BLOCK + threadIdx.x] = rk[q];
            vals[q * BLOCK + threadIdx.x] = rv[q];
          }
          __syncthreads();
        }
      }
      if (!ranked) {
        bitonic_sort_groups<TS, G, BLOCK>(keys, vals, lane, true);
        __syncthreads();
        const long long o = MODE == 3 ? s_base[g] : (active ? R.off[fr] : 0);
        const bool fits = o + cnt <= R.out_limit;
        if (active && !ovf && !fits && lane == 0) R.stats[3] = 1ull;
        if (active && !ovf && fits) {
          for (int s = lane; s < cnt; s += G) {
            __builtin_nontemporal_store(keys[s], &tile.c_col[o + s]);
            __builtin_nontemporal_store((T)vals[s], &tile.c_val[o + s]);
          }
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------------------------
// Sort-based rows (ESC: expand, sort, compress), one wavefront per row, for rows of 64 * K products or fewer.
// Rows of a SUMMA-stage tile product are mostly output-dense (nnz(C_i) close to the products), where a hash table
// costs two LDS slots per output entry plus a rank structure, and caps how many rows a CU holds. Here the products
// of a row are expanded into registers as 32-bit keys (column << 13 | product index), sorted with a bitonic network
// across the wavefront (registers and lane shuffles), and equal columns are summed in product order (the reference's
// A-row order, in float64, rounded once). LDS holds only the float64 products (8 bytes per product) and the sorted
// columns. The count is exact after the sort, so the predicted mode never overflows: a row that outgrows its slot
// is bump-allocated in the same pass.
//   MODE 0 symbolic -> nnz[fr]; MODE 1 numeric at off[fr]; MODE 2 at off[fr] when nnz <= cap[fr], else
//   bump-allocated; MODE 3 bump-allocated.
// ------------------------------------------------------------------------------------------------------------

constexpr int kEscWaves = 4;  // wavefronts per block, each on its own row

constexpr unsigned kEscKeyShift = 32 - kEscColBits;      // key = column << 13 | product index
constexpr unsigned kEscIdxMask = (1u << kEscKeyShift) - 1u;
constexpr unsigned kEscPad = 0xffffffffu;                 // sorts after every real key

// Does sorted position k * 64 + lane start an output entry (a real key whose column differs from its
// predecessor's)? Position 0 always does: the pad key's column field equals the largest column.
template <int K>
__device__ inline bool esc_head(const unsigned (&key)[K], int k, int lane) {
  const unsigned up = (unsigned)__shfl_up((int)key[k], 1, 64);
  const unsigned wrap = k > 0 ? (unsigned)__shfl((int)key[k > 0 ? k - 1 : 0], 63, 64) : kEscPad;
  const unsigned prev = lane == 0 ? wrap : up;
  return key[k] != kEscPad && ((k == 0 && lane == 0) || (prev >> kEscKeyShift) != (key[k] >> kEscKeyShift));
}

// For the K products x = x0 + k * 64 + lane of a wavefront, the staged entry e holding product t = x - base
// (pre[e] <= t < pre[e + 1], pre has ne + 1 prefix counts): K branchless binary searches in lockstep. Products outside
// the chunk get some valid e (the caller skips them).
template <int K>
__device__ inline void esc_lookup(const int* pre, int ne, int base, int lane, int x0, int (&lo)[K]) {
  int top = 1;  // largest power of two < ne (0 when ne == 1)
  while (top < ne) top <<= 1;
  top >>= 1;
#pragma unroll
  for (int k = 0; k < K; k++) lo[k] = 0;
  for (int step = top; step > 0; step >>= 1) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int c = lo[k] + step;
      if (c < ne && pre[c] <= x0 + k * 64 + lane - base) lo[k] = c;
    }
  }
}

// Compare-exchange of one bitonic stage whose partner is `d` lanes away (d < 64): keep min or max of the pair.
// Value of lane (lane ^ D) of a wavefront. Distances 1, 2 and 8 are one DPP move (quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_ror:8), 4 is two DPP moves (row_shl:4 for lanes with bit 2 clear, row_shr:4 for the others), 16 is
// a ds_swizzle in xor mode (within 32 lanes), 32 a bpermute.
template <int D>
__device__ inline unsigned lane_xor(unsigned v, int lane) {
  const int x = (int)v;
  if constexpr (D == 1) return (unsigned)__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
  else if constexpr (D == 2) return (unsigned)__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
  else if constexpr (D == 4) {
    const int up = __builtin_amdgcn_update_dpp(0, x, 0x104, 0xF, 0xF, false);  // lane + 4
    const int dn = __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // lane - 4
    return (unsigned)((lane & 4) ? dn : up);
  } else if constexpr (D == 8) return (unsigned)__builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false);
  else if constexpr (D == 16) return (unsigned)__builtin_amdgcn_ds_swizzle(x, 0x401F);
  else return (unsigned)__shfl_xor(x, D, 64);
}

template <int D>
__device__ inline unsigned esc_xchg(unsigned v, bool keep_min, int lane) {
  const unsigned o = lane_xor<D>(v, lane);
  return keep_min ? min(v, o) : max(v, o);
}

// Bitonic sort, ascending, of the 64 * K keys of a wavefront; element e = k * 64 + lane lives in key[k].
template <int K>
__device__ inline void esc_sort(unsigned (&key)[K], int lane) {
#pragma unroll
  for (int size = 2; size <= 64 * K; size <<= 1) {
#pragma unroll
    for (int d = size >> 1; d > 0; d >>= 1) {
      if (d >= 64) {
        const int dk = d >> 6;
#pragma unroll
        for (int k = 0; k < K; k++) {
          if ((k & dk) == 0) {
            const bool asc = ((k * 64) & size) == 0;
            const unsigned a = key[k], b = key[k | dk];
            const bool sw = (a > b) == asc;
            key[k] = sw ? b : a;
            key[k | dk] = sw ? a : b;
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
          const bool asc = (((k * 64) | lane) & size) == 0;
          const bool lower = (lane & d) == 0;
          const bool keep_min = asc == lower;
          switch (d) {
            case 1: key[k] = esc_xchg<1>(key[k], keep_min, lane); break;
            case 2: key[k] = esc_xchg<2>(key[k], keep_min, lane); break;
            case 4: key[k] = esc_xchg<4>(key[k], keep_min, lane); break;
            case 8: key[k] = esc_xchg<8>(key[k], keep_min, lane); break;
            case 16: key[k] = esc_xchg<16>(key[k], keep_min, lane); break;
            default: key[k] = esc_xchg<32>(key[k], keep_min, lane); break;
          }
        }
      }
    }
  }
}

// Half-cleaner stages at distances 32 * K .. 1 (ascending) over one wavefront's 64 * K keys, which merge a bitonic
// sequence into sorted order: in-register stages for distances of 64 or more, lane exchanges below.
template <int K>
__device__ inline void esc_half_clean(unsigned (&key)[K], int lane) {
#pragma unroll
  for (int d = 32 * K; d > 0; d >>= 1) {
    if (d >= 64) {
      const int dk = d >> 6;
#pragma unroll
      for (int k = 0; k < K; k++) {
        if ((k & dk) == 0) {
          const unsigned a = key[k], b = key[k | dk];
          key[k] = min(a, b);
          key[k | dk] = max(a, b);
        }
      }
    } else {
      const bool keep_min = (lane & d) == 0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        switch (d) {
          case 1: key[k] = esc_xchg<1>(key[k], keep_min, lane); break;
          case 2: key[k] = esc_xchg<2>(key[k], keep_min, lane); break;
          case 4: key[k] = esc_xchg<4>(key[k], keep_min, lane); break;
          case 8: key[k] = esc_xchg<8>(key[k], keep_min, lane); break;
          case 16: key[k] = esc_xchg<16>(key[k], keep_min, lane); break;
          default: key[k] = esc_xchg<32>(key[k], keep_min, lane); break;
        }
      }
    }
  }
}

template <int K, int MODE, typename T, typename RP>
__global__ __launch_bounds__(64 * kEscWaves) void k_esc(RowsDev<T, RP> R, int bin) {
  constexpr int N = 64 * K;
  constexpr bool NUM = MODE != 0;
  __shared__ double s_val[kEscWaves][NUM ? N : 1];  // float64 products by product index
  __shared__ unsigned s_key[kEscWaves][NUM ? N : 1];  // sorted keys
  __shared__ int s_pre[kEscWaves][65];              // per chunk of A entries: products before each entry
  __shared__ RP s_st[kEscWaves][64];                // B-row start of each entry
  __shared__ T s_av[kEscWaves][64];                 // A value of each entry
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int count = *R.count_of(bin);
  const int* items = R.bin_items + (int64_t)bin * R.bin_stride;
  for (int idx = blockIdx.x * kEscWaves + w; idx < count; idx += gridDim.x * kEscWaves) {
    const int fr = items[idx];
    const TileDev<T, RP> tile = R.tiles[R.blk_tile[fr >> 8]];
    const int64_t i = fr - tile.flat0;
    const RP a0 = tile.a_rp[i], a1 = tile.a_rp[i + 1];
    unsigned key[K];
#pragma unroll
    for (int k = 0; k < K; k++) key[k] = kEscPad;
    // expand: product x (A-row order) -> key (column << 13 | x), float64 product -> s_val[x]
    int base = 0;
    for (RP c0 = a0; c0 < a1; c0 += 64) {
      const int ne = (int)((a1 - c0) < (RP)64 ? (a1 - c0) : (RP)64);
      int len = 0;
      if (lane < ne) {
        const int kk = tile.a_col[c0 + lane];
        const RP st = tile.b_rp[kk];
        len = (int)(tile.b_rp[kk + 1] - st);
        s_st[w][lane] = st;
        if (NUM) s_av[w][lane] = tile.a_val[c0 + lane];
      }
      int inc = len;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
      }
      const int total = __shfl(inc, 63, 64);
      if (lane < ne) s_pre[w][lane] = inc - len;
      if (lane == 0) s_pre[w][ne] = total;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      int lo[K];  // entry e with s_pre[e] <= t < s_pre[e + 1], all K lookups in lockstep
      esc_lookup<K>(s_pre[w], ne, base, lane, 0, lo);
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int x = k * 64 + lane;
        const int t = x - base;
        if (t >= 0 && t < total) {
          const RP q = s_st[w][lo[k]] + (RP)(t - s_pre[w][lo[k]]);
          key[k] = ((unsigned)tile.b_col[q] << kEscKeyShift) | (unsigned)x;
          if (NUM) s_val[w][x] = (double)s_av[w][lo[k]] * (double)tile.b_val[q];
        }
      }
      base += total;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    esc_sort<K>(key, lane);
    // compress: a sorted position starts an output entry when its column differs from the previous position's
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < K; k++) cnt += __popcll(__ballot(esc_head<K>(key, k, lane)));
    if (MODE == 0) {
      if (lane == 0) R.nnz[fr] = cnt;
      continue;
    }
    long long o = 0;
    if (lane == 0) {
      if (MODE == 1 || (MODE == 2 && cnt <= R.cap[fr])) {
        o = R.off[fr];
      } else {
        o = (long long)atomicAdd(R.cursor, (unsigned long long)cnt);
        R.off[fr] = o;
      }
      if (o + cnt > R.out_limit) {  // refused: nothing is written and, after this pass, the row holds 0 entries
        R.stats[3] = 1ull;
        o = -1;
        if (MODE != 1) R.off[fr] = 0;
      }
      R.nnz[fr] = (o < 0 && MODE != 1) ? 0 : cnt;
    }
    o = __shfl(o, 0, 64);
#pragma unroll
    for (int k = 0; k < K; k++) s_key[w][k * 64 + lane] = key[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (o >= 0) {
      int before = 0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const bool head = esc_head<K>(key, k, lane);
        const unsigned long long mask = __ballot(head);
        if (head) {
          // sum the run of equal columns in sorted order, i.e. in product (A-row) order
          const unsigned col = key[k] >> kEscKeyShift;
          double sum = s_val[w][key[k] & kEscIdxMask];
          for (int p = k * 64 + lane + 1; p < N; p++) {
            const unsigned kp = s_key[w][p];
            if (kp == kEscPad || (kp >> kEscKeyShift) != col) break;
            sum += s_val[w][kp & kEscIdxMask];
          }
          const long long pos = o + before + __popcll(mask & ((1ull << lane) - 1ull));
          __builtin_nontemporal_store((int)col, &tile.c_col[pos]);
          __builtin_nontemporal_store((T)sum, &tile.c_val[pos]);
        }
        before += __popcll(mask);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// ------------------------------------------------------------------------------------------------------------
// Rows of at most 32 products, half a wavefront each (two rows per wavefront). The products are expanded one per lane
// as 32-bit keys (column << 5 | product index), sorted by a bitonic network over the 32 lanes (DPP and swizzle
// exchanges: no LDS, no barrier), and equal columns are summed in product (A-row) order, in float64, rounded once.
// A hash row of this size costs a table clear, LDS atomics and a sort of its table in LDS; a one-wavefront ESC row
// leaves half the lanes idle. Tiles up to 2^27 - 2 columns wide. Modes as in k_esc.
// ------------------------------------------------------------------------------------------------------------

constexpr int kSmallWaves = 4;  // wavefronts per block

// One bitonic stage (block size K, partner distance J) over the 32 elements of a row group: element x = s * G + gl
// sits in slot s of lane gl. Partners J >= G lie in the same lane (another slot); closer ones in lane gl ^ J.
template <int K, int J, int G, int S>
__device__ inline void grp_cx(unsigned (&key)[S], int gl) {
  if constexpr (J >= G) {
    constexpr int D = J / G;
#pragma unroll
    for (int s = 0; s < S; s++)
      if ((s & D) == 0) {
        const bool asc = ((s * G + gl) & K) == 0;
        const unsigned a = key[s], b = key[s + D];
        key[s] = asc ? min(a, b) : max(a, b);
        key[s + D] = asc ? max(a, b) : min(a, b);
      }
  } else {
#pragma unroll
    for (int s = 0; s < S; s++) {
      const unsigned other = lane_xor<J>(key[s], gl);
      const int x = s * G + gl;
      const bool lower = (x & J) == 0, asc = (x & K) == 0;
      key[s] = lower == asc ? min(key[s], other) : max(key[s], other);
    }
  }
}

// ascending over the N elements of each row group (bitonic: block sizes K = 2 .. N, partner distances J = K/2 .. 1)
template <int K, int J, int G, int S, int N>
__device__ inline void grp_sort_from(unsigned (&key)[S], int gl) {
  grp_cx<K, J, G, S>(key, gl);
  if constexpr (J > 1) grp_sort_from<K, J / 2, G, S, N>(key, gl);
  else if constexpr (K < N) grp_sort_from<K * 2, K, G, S, N>(key, gl);
}
template <int G, int S, int N>
__device__ inline void grp_sort(unsigned (&key)[S], int gl) {
  grp_sort_from<2, 1, G, S, N>(key, gl);
}

// The value of slot-array v at element x of the group (lane x % G, slot x / G). Every lane runs the exchanges.
template <int G, int S, typename V>
__device__ inline V grp_at(const V (&v)[S], int x) {
  V r = V(0);
#pragma unroll
  for (int s = 0; s < S; s++) {
    const V t = __shfl(v[s], x & (G - 1), G);
    if (x / G == s) r = t;
  }
  return r;
}

// One row per group of G lanes, 32 / G products per lane (G = 8: eight rows per wavefront; measured on the FP64
// Laplacian square, 16.7 M rows of 25 products: G = 32 9.4 ms, 16 7.5 ms, 8 6.6 ms, 4 8.5 ms): expand, sort,
// compress, write. `next` runs once the row's last global loads are issued (the caller's prefetch of its next row
// overlaps them). Every exchange is executed by all groups of the wavefront.
template <int MODE, int G, int N, typename T, typename RP, typename NEXT>
__device__ inline void small_row(const RowsDev<T, RP>& R, const TileDev<T, RP>& tile, int fr, bool active, RP a0,
                                 RP a1, int grp, int gl, NEXT&& next) {
  constexpr bool NUM = MODE != 0;
  constexpr int S = N / G;
  constexpr unsigned kShift = N == 32 ? 5 : 6;  // product-index bits of a key
  static_assert(N == 32 || N == 64, "rows of at most 32 or 64 products");
  const int na = (int)(a1 - a0);
  int na_max = na;
#pragma unroll
  for (int d = G; d < 64; d <<= 1) na_max = max(na_max, __shfl_xor(na_max, d, 64));
  // expand: product x (A-row order) goes to slot x / G of lane x % G of the group
  unsigned key[S];
  double pv[S];
#pragma unroll
  for (int q = 0; q < S; q++) {
    key[q] = kEscPad;
    pv[q] = 0.0;
  }
  int base = 0;
  for (int c0 = 0; c0 < na_max; c0 += G) {
    const int ne = min(max(na - c0, 0), G);
    int len = 0;
    RP st = 0;
    T av = T(0);
    if (gl < ne) {
      const int kk = tile.a_col[a0 + c0 + gl];
      st = tile.b_rp[kk];
      len = (int)(tile.b_rp[kk + 1] - st);
      if (NUM) av = tile.a_val[a0 + c0 + gl];
    }
    int inc = len;
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
      const int t = __shfl_up(inc, d, G);
      if (gl >= d) inc += t;
    }
    const int ex = inc - len;
    const int total = __shfl(inc, G - 1, G);
#pragma unroll
    for (int q = 0; q < S; q++) {
      const int t = gl + G * q - base;
      int lo = 0;  // the entry holding product t: the last e < ne with ex[e] <= t
#pragma unroll
      for (int step = G / 2; step > 0; step >>= 1) {
        const int c = lo + step;
        const int v = __shfl(ex, c < G ? c : G - 1, G);
        if (c < ne && v <= t) lo = c;
      }
      const RP st_e = __shfl(st, lo, G);
      const int ex_e = __shfl(ex, lo, G);
      const T av_e = __shfl(av, lo, G);
      if (t >= 0 && t < total) {  // the row has at most N products, so base + total <= N
        const RP qb = st_e + (RP)(t - ex_e);
        key[q] = ((unsigned)tile.b_col[qb] << kShift) | (unsigned)(gl + G * q);
        if (NUM) pv[q] = (double)av_e * (double)tile.b_val[qb];
      }
    }
    base += total;
  }
  next();
  grp_sort<G, S, N>(key, gl);
  // compress: a sorted position starts an output entry when its column differs from the previous position's
  unsigned hm[S];
  bool head[S];
  int cnt = 0;
#pragma unroll
  for (int q = 0; q < S; q++) {
    const unsigned up = __shfl_up(key[q], 1, G);
    const unsigned wrap = q > 0 ? __shfl(key[q > 0 ? q - 1 : 0], G - 1, G) : 0u;
    const unsigned prev = gl > 0 ? up : wrap;
    head[q] = key[q] != kEscPad && (gl + q * G == 0 || (prev >> kShift) != (key[q] >> kShift));
    hm[q] = (unsigned)((__ballot(head[q]) >> (grp * G)) & (G == 32 ? 0xffffffffull : ((1ull << G) - 1)));
    cnt += __popc(hm[q]);
  }
  if (MODE == 0) {
    if (active && gl == 0) R.nnz[fr] = cnt;
    return;
  }
  long long o = 0;
  if (active && gl == 0) {
    if (MODE == 1 || (MODE == 2 && cnt <= R.cap[fr])) {
      o = R.off[fr];
    } else {
      o = (long long)atomicAdd(R.cursor, (unsigned long long)cnt);
      R.off[fr] = o;
    }
    if (o + cnt > R.out_limit) {  // refused: nothing is written and, after this pass, the row holds 0 entries
      R.stats[3] = 1ull;
      o = -1;
      if (MODE != 1) R.off[fr] = 0;
    }
    R.nnz[fr] = (o < 0 && MODE != 1) ? 0 : cnt;
  }
  o = __shfl(o, 0, G);
  // sum each run of equal columns in sorted order, i.e. in product order, by the run's first element
  double mine[S], sum[S];
  bool run[S];
#pragma unroll
  for (int q = 0; q < S; q++) {
    const int pidx = (int)(key[q] & (unsigned)(N - 1));
    double got = 0.0;
#pragma unroll
    for (int r = 0; r < S; r++) {
      const double t = __shfl(pv[r], pidx & (G - 1), G);
      if (pidx / G == r) got = t;
    }
    mine[q] = got;
    sum[q] = got;
    run[q] = head[q];
  }
  for (int j = 1; j < N; j++) {
    bool any = false;
#pragma unroll
    for (int q = 0; q < S; q++) {
      const int x = q * G + gl + j;  // the element j places after this one
      const unsigned kj = grp_at<G, S>(key, x);
      const double vj = grp_at<G, S>(mine, x);
      run[q] = run[q] && x < N && kj != kEscPad && (kj >> kShift) == (key[q] >> kShift);
      if (run[q]) sum[q] += vj;
      any = any || run[q];
    }
    if (__ballot(any) == 0ull) break;
  }
  if (o >= 0) {
    int before = 0;
#pragma unroll
    for (int q = 0; q < S; q++) {
      if (head[q]) {
        const long long pos = o + before + __popc(hm[q] & ((1u << gl) - 1u));
        // plain stores here: rows this small write a fraction of a cache line each, and L2 merges neighbouring
        // rows' pieces before they reach HBM (streaming stores made a 4096^2-grid Laplacian 6.6 -> 7.2 ms)
        tile.c_col[pos] = (int)(key[q] >> kShift);
        tile.c_val[pos] = (T)sum[q];
      }
      before += __popc(hm[q]);
    }
  }
}

// ONE: every row belongs to tiles[0] (a single product): the tile is read once, and the next row's bin item and
// A-row bounds are loaded while the current row runs, so a row waits on three dependent loads (A columns, B-row
// bounds, B entries) instead of five. G: lanes per row (32 or 16).
template <int MODE, bool ONE, int G, int N, typename T, typename RP>
__global__ __launch_bounds__(64 * kSmallWaves) void k_esc_small(RowsDev<T, RP> R, int bin) {
  constexpr int RW = 64 / G;  // rows per wavefront
  const int lane = threadIdx.x & 63, grp = lane / G, gl = lane % G;
  const int count = *R.count_of(bin);
  const int* items = R.bin_items + (int64_t)bin * R.bin_stride;
  const int stride = gridDim.x * kSmallWaves * RW;
  int64_t p = (int64_t)(blockIdx.x * kSmallWaves + (threadIdx.x >> 6)) * RW;  // the loop is wave-uniform
  if constexpr (ONE) {
    const TileDev<T, RP> tile = R.tiles[0];
    bool active = p + grp < count;
    int fr = active ? items[p + grp] : 0;
    RP a0 = 0, a1 = 0;
    if (active) {
      a0 = tile.a_rp[fr - tile.flat0];
      a1 = tile.a_rp[fr - tile.flat0 + 1];
    }
    for (; p < count; p += stride) {
      const int64_t idx_n = p + stride + grp;
      const bool act_n = idx_n < count;
      const int fr_n = act_n ? items[idx_n] : 0;
      RP a0_n = 0, a1_n = 0;
      small_row<MODE, G, N>(R, tile, fr, active, a0, a1, grp, gl, [&] {
        if (act_n) {
          a0_n = tile.a_rp[fr_n - tile.flat0];
          a1_n = tile.a_rp[fr_n - tile.flat0 + 1];
        }
      });
      active = act_n;
      fr = fr_n;
      a0 = a0_n;
      a1 = a1_n;
    }
  } else {
    for (; p < count; p += stride) {
      const int64_t idx = p + grp;
      const bool active = idx < count;
      const int fr = active ? items[idx] : 0;
      TileDev<T, RP> tile;
      RP a0 = 0, a1 = 0;
      if (active) {
        tile = R.tiles[R.blk_tile[fr >> 8]];
        const int64_t i = fr - tile.flat0;
        a0 = tile.a_rp[i];
        a1 = tile.a_rp[i + 1];
      }
      small_row<MODE, G, N>(R, tile, fr, active, a0, a1, grp, gl, [] {});
    }
  }
}

// Sort-based rows of 512 * W products or fewer, W wavefronts per row: each wavefront expands and sorts 512 of the
// row's products in registers (as the single-wavefront kernel does), the sorted runs are merged pairwise in place in
// LDS (each key's merged position is its own index plus its rank in the partner run, found by binary search while
// every key is still in place; ties go to the left run, so the pad keys fill the tail), and the block compresses the
// sorted row. LDS: 12 bytes per product (key and float64 product) plus one staged chunk of A entries.
constexpr int kEscChunk = 512;

template <int W, int MODE, typename T, typename RP>
__global__ __launch_bounds__(64 * W, 6) void k_esc_block(RowsDev<T, RP> R, int bin) {
  constexpr int N = kEscChunk * W;
  constexpr int K = kEscChunk / 64;
  constexpr bool NUM = MODE != 0;
  __shared__ unsigned s_key[N];
  __shared__ double s_val[NUM ? N : 1];
  __shared__ int s_pre[65];
  __shared__ RP s_st[64];
  __shared__ T s_av[64];
  __shared__ int s_cnt[W];
  __shared__ long long s_o;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int count = *R.count_of(bin);
  const int* items = R.bin_items + (int64_t)bin * R.bin_stride;
  for (int idx = blockIdx.x; idx < count; idx += gridDim.x) {
    const int fr = items[idx];
    const TileDev<T, RP> tile = R.tiles[R.blk_tile[fr >> 8]];
    const int64_t i = fr - tile.flat0;
    const RP a0 = tile.a_rp[i], a1 = tile.a_rp[i + 1];
    unsigned key[K];
#pragma unroll
    for (int k = 0; k < K; k++) key[k] = kEscPad;
    // expand: wavefront 0 stages each chunk of 64 A entries; wavefront w takes products x in [w * 512, (w + 1) * 512)
    int base = 0;
    for (RP c0 = a0; c0 < a1; c0 += 64) {
      const int ne = (int)((a1 - c0) < (RP)64 ? (a1 - c0) : (RP)64);
      if (w == 0) {
        int len = 0;
        if (lane < ne) {
          const int kk = tile.a_col[c0 + lane];
          const RP st = tile.b_rp[kk];
          len = (int)(tile.b_rp[kk + 1] - st);
          s_st[lane] = st;
          if (NUM) s_av[lane] = tile.a_val[c0 + lane];
        }
        int inc = len;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const int t = __shfl_up(inc, d, 64);
          if (lane >= d) inc += t;
        }
        if (lane < ne) s_pre[lane] = inc - len;
        if (lane == 63) s_pre[ne] = inc;
      }
      __syncthreads();
      const int total = s_pre[ne];
      int lo[K];
      esc_lookup<K>(s_pre, ne, base, lane, w * kEscChunk, lo);
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int x = w * kEscChunk + k * 64 + lane;
        const int t = x - base;
        if (t >= 0 && t < total) {
          const RP q = s_st[lo[k]] + (RP)(t - s_pre[lo[k]]);
          key[k] = ((unsigned)tile.b_col[q] << kEscKeyShift) | (unsigned)x;
          if (NUM) s_val[x] = (double)s_av[lo[k]] * (double)tile.b_val[q];
        }
      }
      base += total;
      __syncthreads();
    }
    esc_sort<K>(key, lane);
    // merge the wavefronts' sorted runs with the rest of a bitonic network: for each block size S = 1024 .. N, a
    // mirror stage (position E against E ^ (S - 1)) and half-cleaners at distances S/4 .. 1. Distances of 512 or
    // more cross wavefronts and go through LDS; shorter ones stay in registers and lanes.
    const int E0 = w * kEscChunk + lane;  // position of key[k] is E0 + k * 64
    // one compare-exchange stage across wavefronts: partner position E ^ d through LDS; mirror stages (d = S - 1)
    // keep the minimum in the lower half of each S-block, half-cleaners in the lower half of each 2d-block
    auto lds_stage = [&](int d, int S) {
#pragma unroll
      for (int k = 0; k < K; k++) s_key[E0 + k * 64] = key[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int E = E0 + k * 64;
        const unsigned o = s_key[E ^ d];
        const bool lower = d == S - 1 ? (E & (S >> 1)) == 0 : (E & d) == 0;
        key[k] = lower ? min(key[k], o) : max(key[k], o);
      }
      __syncthreads();
    };
#pragma unroll
    for (int S = 2 * kEscChunk; S <= N; S <<= 1) {
      lds_stage(S - 1, S);
      for (int d = S >> 2; d >= kEscChunk; d >>= 1) lds_stage(d, S);
      esc_half_clean<K>(key, lane);
    }
#pragma unroll
    for (int k = 0; k < K; k++) s_key[E0 + k * 64] = key[k];
    __syncthreads();
    // compress: wavefront w owns sorted positions [w * 512, (w + 1) * 512), as k * 64 + lane
    int wcnt = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int p = w * kEscChunk + k * 64 + lane;
      const unsigned prev = p > 0 ? s_key[p - 1] : kEscPad;
      const bool head = key[k] != kEscPad && (p == 0 || (prev >> kEscKeyShift) != (key[k] >> kEscKeyShift));
      wcnt += __popcll(__ballot(head));
    }
    if (lane == 0) s_cnt[w] = wcnt;
    __syncthreads();
    int before = 0, cnt = 0;
#pragma unroll
    for (int q = 0; q < W; q++) {
      if (q < w) before += s_cnt[q];
      cnt += s_cnt[q];
    }
    if (MODE == 0) {
      if (threadIdx.x == 0) R.nnz[fr] = cnt;
      __syncthreads();
      continue;
    }
    if (threadIdx.x == 0) {
      long long o;
      if (MODE == 1 || (MODE == 2 && cnt <= R.cap[fr])) {
        o = R.off[fr];
      } else {
        o = (long long)atomicAdd(R.cursor, (unsigned long long)cnt);
        R.off[fr] = o;
      }
      if (o + cnt > R.out_limit) {  // refused: nothing is written and, after this pass, the row holds 0 entries
        R.stats[3] = 1ull;
        o = -1;
        if (MODE != 1) R.off[fr] = 0;
      }
      R.nnz[fr] = (o < 0 && MODE != 1) ? 0 : cnt;
      s_o = o;
    }
    __syncthreads();
    const long long o = s_o;
    if (o >= 0) {
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int p = w * kEscChunk + k * 64 + lane;
        const unsigned prev = p > 0 ? s_key[p - 1] : kEscPad;
        const bool head = key[k] != kEscPad && (p == 0 || (prev >> kEscKeyShift) != (key[k] >> kEscKeyShift));
        const unsigned long long mask = __ballot(head);
        if (head) {
          const unsigned col = key[k] >> kEscKeyShift;
          double sum = s_val[key[k] & kEscIdxMask];
          for (int p2 = p + 1; p2 < N; p2++) {
            const unsigned kp = s_key[p2];
            if (kp == kEscPad || (kp >> kEscKeyShift) != col) break;
            sum += s_val[kp & kEscIdxMask];
          }
          const long long pos = o + before + __popcll(mask & ((1ull << lane) - 1ull));
          __builtin_nontemporal_store((int)col, &tile.c_col[pos]);
          __builtin_nontemporal_store((T)sum, &tile.c_val[pos]);
        }
        before += __popcll(mask);
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------------------------
// Heavy rows (more products than the sort-based bins take and more output entries than the largest hash table),
// rank-slice accumulator: one workgroup per row, everything in LDS.
//   pass 1: LDS bitmap of the row's columns (one atomicOr per product) and prefix counts per group of kHRGroup
//           words, so every column's output rank is known and nnz(C_i) is exact; the columns are written straight
//           from the bitmap;
//   pass 2 (numeric modes), per slice of hr_slice() ranks: the products whose columns fall in the slice's column range
//           are added in float64 into an LDS array indexed by rank - slice start (ds_add_f64), then written out with
//           unit stride. Most rows fit one slice; larger rows enumerate their products once per slice.
// Modes as in k_hash: 0 symbolic, 1 numeric at off[fr], 2 predicted (slot if nnz <= cap, else bump-allocate),
// 3 bump-allocate. Needs n <= 32 * kHRWords; wider tiles use k_heavy (global dense accumulator).
// ------------------------------------------------------------------------------------------------------------

constexpr int kHRWords = 16384;  // bitmap words: tiles up to 2^19 columns
// Words per prefix group: a column's rank is its group's prefix plus the bits before it in the group, read as one
// 16-byte LDS load. Narrow groups cost prefix space (16 KB for 2^19 columns) but keep the per-product LDS reads at
// 20 bytes; 16-word groups (four loads) were measured 0.75 ms slower on the s22 stage despite larger slices.
constexpr int kHRGroup = 4;
// float64 accumulators per slice: what LDS holds beside the bitmap, the prefixes and the entry staging (64 KB or 56 KB)
template <typename T, typename RP>
constexpr int hr_slice() { return sizeof(T) == 4 && sizeof(RP) == 4 ? 8192 : 7168; }
// symbolic: 512 threads, two workgroups per CU; numeric: 1024 threads, one workgroup per CU (bitmap + slice)
template <int MODE>
constexpr int hr_block() { return MODE == 0 ? 512 : 1024; }

typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
static_assert(kHRGroup == 4, "hr_rank reads one 16-byte group");

// Output rank of column j: set bits before j. The group prefix counts the bits before j's group; the group (j's word
// and the words before it) is one 16-byte LDS load.
__device__ inline int hr_rank(const unsigned* bits, const int* gpre, int j) {
  const int w = j >> 5, g = w / kHRGroup, wi = w % kHRGroup;
  const u32x4_t x = *reinterpret_cast<const u32x4_t*>(bits + g * kHRGroup);
  const unsigned below = (1u << (j & 31)) - 1u;
  return gpre[g] + (wi > 0 ? __popc(x.x) : 0) + (wi > 1 ? __popc(x.y) : 0) + (wi > 2 ? __popc(x.z) : 0) +
         __popc((wi == 0 ? x.x : wi == 1 ? x.y : wi == 2 ? x.z : x.w) & below);
}

// Column of output rank r (0 <= r < number of set bits): the r-th set bit. One thread.
__device__ inline int hr_select(const unsigned* bits, const int* gpre, int ng, int r) {
  int lo = 0, hi = ng;  // last group with gpre <= r
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (gpre[mid] <= r) lo = mid; else hi = mid;
  }
  int left = r - gpre[lo];
  int w = lo * kHRGroup;
  while (true) {
    const int c = __popc(bits[w]);
    if (left < c) break;
    left -= c;
    w++;
  }
  unsigned word = bits[w];
  for (int q = 0; q < left; q++) word &= word - 1;
  return w * 32 + (__ffs(word) - 1);
}

// First position in [lo0, len) of the B row at st whose column is >= c (binary search in global memory).
template <typename RP>
__device__ inline RP hr_lower_bound(const int* b_col, RP st, RP lo0, RP len, int c) {
  RP lo = lo0, hi = len;
  while (lo < hi) {
    const RP mid = (lo + hi) >> 1;
    if (b_col[st + mid] < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Orders the heavy bin by decreasing products (classes of equal floor(log2 P)), so that the dynamic queue of
// k_heavy_rank starts the largest rows first and the small ones fill in behind them. One block.
template <typename T, typename RP>
__global__ __launch_bounds__(1024) void k_order_heavy(RowsDev<T, RP> R) {
  __shared__ int s_hist[32], s_next[32];
  const int cnt = *R.count_of(kHeavyBin);
  int* items = R.bin_items + (int64_t)kHeavyBin * R.bin_stride;
  if (threadIdx.x < 32) s_hist[threadIdx.x] = 0;
  __syncthreads();
  auto cls = [&](int fr) {  // 31 - floor(log2 P), clamped to [0, 31]: larger rows first
    const long long p = R.prod[fr];
    const int lg = 63 - __clzll((unsigned long long)(p > 1 ? p : 1));
    return 31 - (lg < 31 ? lg : 31);
  };
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) atomicAdd(&s_hist[cls(items[q])], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int c = 0; c < 32; c++) {
      s_next[c] = run;
      run += s_hist[c];
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    const int fr = items[q];
    R.tmp_items[atomicAdd(&s_next[cls(fr)], 1)] = fr;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) items[q] = R.tmp_items[q];
}

template <int MODE, typename T, typename RP>
__global__ __launch_bounds__(hr_block<MODE>()) void k_heavy_rank(RowsDev<T, RP> R, int bin) {
  constexpr int BLOCK = hr_block<MODE>();
  constexpr bool NUM = MODE != 0;
  constexpr int kHRSlice = hr_slice<T, RP>();
  __shared__ __align__(16) unsigned s_bits[kHRWords];
  __shared__ int s_gpre[NUM ? kHRWords / kHRGroup : 1];  // the symbolic pass needs only the total
  __shared__ __align__(16) double s_acc[NUM ? kHRSlice : 1];  // a slice's columns (as int), then its values
  __shared__ long long s_scan[BLOCK / 64 + 1];
  __shared__ long long s_base;
  __shared__ int s_crange[2];
  __shared__ RP s_est[BLOCK];
  // product prefixes of one staged chunk fit 32 bits: at most BLOCK entries, each a B row of at most 2^19 columns
  __shared__ int s_epre[BLOCK + 1];
  __shared__ T s_eval[BLOCK];
  __shared__ long long s_lscan[BLOCK / 64 + 1];
  constexpr int kGroupsPerThread = (kHRWords / kHRGroup + BLOCK - 1) / BLOCK;
  const int tid = threadIdx.x;
  const int count = *R.count_of(bin);
  const int* items = R.bin_items + (int64_t)bin * R.bin_stride;
  int* s_colbuf = reinterpret_cast<int*>(s_acc);
  __shared__ int s_idx;
  // Rows are taken from a queue (largest first), so one block's share is not fixed in advance. The next row is
  // fetched while the current one runs: its queue slot when the current row starts, then one dependent load per
  // barrier of the current row (bin item, tile, A row bounds, each entry's A column and value, its B-row range). A
  // row then starts with its staging in registers instead of behind a chain of global round trips.
  int nq = 0;        // thread 0: the next row's queue slot
  int pst = 0;       // fetch steps done for the next row (uniform)
  int n_idx = 0, n_fr = 0, n_tix = 0, n_kk = 0, n_cap = 0;  // n_cap, n_off: thread 0 only
  long long n_off = 0;
  RP n_a0 = 0, n_a1 = 0, n_st = 0, n_len = 0;
  T n_av = T(0);
  auto advance = [&](int upto) {  // all threads, uniform upto; a barrier separates steps 1 and 2
    if (pst < 1 && upto >= 1) {
      if (tid == 0) s_idx = nq;
      pst = 1;
      if (upto >= 2) __syncthreads();
    }
    if (pst < 2 && upto >= 2) {
      n_idx = __builtin_amdgcn_readfirstlane(s_idx);
      if (n_idx < count) n_fr = items[n_idx];
      pst = 2;
    }
    if (pst < 3 && upto >= 3) {
      if (n_idx < count) n_tix = R.blk_tile[n_fr >> 8];
      pst = 3;
    }
    if (pst < 4 && upto >= 4) {
      if (n_idx < count) {
        const TileDev<T, RP>& t = R.tiles[n_tix];
        const int64_t ni = n_fr - t.flat0;
        n_a0 = t.a_rp[ni];
        n_a1 = t.a_rp[ni + 1];
        if (NUM && tid == 0) {  // only this block reads or rewrites the row's slot
          n_cap = R.cap[n_fr];
          n_off = R.off[n_fr];
        }
      }
      pst = 4;
    }
    if (pst < 5 && upto >= 5) {
      if (n_idx < count && n_a1 - n_a0 <= (RP)BLOCK && tid < (int)(n_a1 - n_a0)) {
        const TileDev<T, RP>& t = R.tiles[n_tix];
        n_kk = t.a_col[n_a0 + tid];
        if (NUM) n_av = t.a_val[n_a0 + tid];
      }
      pst = 5;
    }
    if (pst < 6 && upto >= 6) {
      if (n_idx < count && n_a1 - n_a0 <= (RP)BLOCK && tid < (int)(n_a1 - n_a0)) {
        const TileDev<T, RP>& t = R.tiles[n_tix];
        n_st = t.b_rp[n_kk];
        n_len = t.b_rp[n_kk + 1] - n_st;
      }
      pst = 6;
    }
  };
  if (tid == 0) nq = (int)atomicAdd(R.heavy_queue, 1ull);
  advance(6);  // the first row: the whole chain, unhidden
  for (;;) {
    if (n_idx >= count) break;
    const int fr = n_fr;
    const TileDev<T, RP> tile = R.tiles[n_tix];
    const int64_t i = fr - tile.flat0;
    const int W = (tile.n + 31) / 32;
    const int Wr = (W + kHRGroup - 1) / kHRGroup * kHRGroup;  // whole groups (the tail words stay zero)
    for (int w = tid; w < Wr; w += BLOCK) s_bits[w] = 0u;
    const RP a0 = n_a0, a1 = n_a1;
    const int na = (int)(a1 - a0);
    const bool one_chunk = na <= BLOCK;  // one entry per thread, staged once and kept in registers
    const RP my_st = n_st, my_len = n_len;
    const T my_av = n_av;
    const int my_cap = n_cap;
    const long long my_off = n_off;
    if (tid == 0) nq = (int)atomicAdd(R.heavy_queue, 1ull);
    pst = 0;
    n_st = 0;
    n_len = 0;
    n_av = T(0);
    do {
      // stage (start, length) of every entry from registers and run the product loop
      auto run_staged = [&](RP st, RP len, auto&& f, auto num_tag) {
        constexpr bool kNum = decltype(num_tag)::value;
        if (tid < na) {
          s_est[tid] = st;
          if (kNum) s_eval[tid] = my_av;
        }
        long long tot;
        const long long ex = block_excl_scan<BLOCK, long long>((long long)(tid < na ? len : 0), &tot, s_lscan);
        if (tid < na) s_epre[tid] = (int)ex;
        if (tid == 0) s_epre[na] = (int)tot;
        __syncthreads();
        enumerate_staged<BLOCK, kNum, 4>(tile, na, tot, tid, s_est, s_epre, s_eval, nullptr, f);
        __syncthreads();
      };
      __syncthreads();
      auto set_bit = [&](int j, double) { atomicOr(&s_bits[j >> 5], 1u << (j & 31)); };
      if (one_chunk) run_staged(my_st, my_len, set_bit, std::false_type{});
      else for_each_product<BLOCK, BLOCK, false, T, RP>(tile, a0, a1, tid, s_est, s_epre, s_eval, s_lscan, nullptr,
                                                        set_bit);
      advance(1);
      __syncthreads();
      advance(2);
      // group prefixes, in order: thread t counts groups [t * kGroupsPerThread, (t + 1) * kGroupsPerThread)
      const int ng = Wr / kHRGroup;
      int gc[kGroupsPerThread];
      long long c = 0;
#pragma unroll
      for (int q = 0; q < kGroupsPerThread; q++) {
        const int g = tid * kGroupsPerThread + q;
        int n1 = 0;
        if (g < ng) {
          const u32x4_t* p4 = reinterpret_cast<const u32x4_t*>(s_bits + g * kHRGroup);
#pragma unroll
          for (int v = 0; v < kHRGroup / 4; v++) {
            const u32x4_t x = p4[v];
            n1 += __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
          }
        }
        gc[q] = n1;
        c += n1;
      }
      long long total;
      long long run = block_excl_scan<BLOCK, long long>(c, &total, s_scan);
      advance(3);
      const int run0 = (int)run;
#pragma unroll
      for (int q = 0; q < kGroupsPerThread; q++) {
        const int g = tid * kGroupsPerThread + q;
        if (NUM && g < ng) s_gpre[g] = (int)run;
        run += gc[q];
      }
      // A row whose columns fit one slice (most heavy rows) stages them here, by rank, from the words this thread
      // just counted: no separate pass over the bitmap and no column search.
      if (NUM && total <= (long long)kHRSlice) {
        int r = run0;
        for (int w = tid * kGroupsPerThread * kHRGroup; w < (tid + 1) * kGroupsPerThread * kHRGroup && w < Wr; w++) {
          unsigned m = s_bits[w];
          while (m) {
            s_colbuf[r++] = w * 32 + (__ffs(m) - 1);
            m &= m - 1;
          }
        }
      }
      if (MODE == 0) {
        if (tid == 0) {
          R.nnz[fr] = (int)total;
          atomicAdd(&R.stats[1], 1ull);
        }
        __syncthreads();
        break;
      }
      if (tid == 0) {
        long long b;
        if (MODE == 1 || (MODE == 2 && total <= (long long)my_cap)) {
          b = my_off;
        } else {
          b = (long long)atomicAdd(R.cursor, (unsigned long long)total);
          R.off[fr] = b;
        }
        if (b + total > R.out_limit) {  // refused: nothing is written and, after this pass, the row holds 0 entries
          R.stats[3] = 1ull;
          b = -1;
          if (MODE != 1) R.off[fr] = 0;
        }
        s_base = b;
        R.nnz[fr] = (b < 0 && MODE != 1) ? 0 : (int)total;
        atomicAdd(&R.stats[1], 1ull);
      }
      __syncthreads();
      const long long base = s_base;
      advance(4);
      if (base < 0) {  // refused: nothing is written
        __syncthreads();
        break;
      }
      // one rank slice at a time: columns (staged in LDS by rank, then stored with unit stride), then values. A
      // slice's first column and each entry's first B position in it are where the previous slice ended.
      RP my_lo = 0;
      for (long long s0 = 0; s0 < total; s0 += kHRSlice) {
        const int len = (int)(total - s0 < kHRSlice ? total - s0 : kHRSlice);
        const bool whole = s0 == 0 && len == total;  // the columns are staged already (above)
        int c_lo = 0, c_hi = INT_MAX;
        if (!whole) {
          if (tid == 0) {  // column range of the slice
            s_crange[0] = s0 == 0 ? 0 : s_crange[1];
            s_crange[1] = s0 + len >= total ? INT_MAX : hr_select(s_bits, s_gpre, ng, (int)(s0 + len));
          }
          __syncthreads();
          c_lo = s_crange[0];
          c_hi = s_crange[1];
        }
        const int w_lo = c_lo >> 5, w_hi = c_hi == INT_MAX ? W - 1 : (c_hi - 1) >> 5;
        // a wavefront takes 64 consecutive words (whole groups) at a time: each word's rank base is its group prefix
        // plus the bits of the group's earlier words (a kHRGroup-lane scan of popcounts)
        if (!whole)
        for (int wb = (w_lo & ~63) + (tid & ~63); wb <= w_hi; wb += BLOCK) {
          const int lane = tid & 63, w = wb + lane;
          const unsigned full = w < W ? s_bits[w] : 0u;
          const int pc = __popc(full);
          int inc = pc;
#pragma unroll
          for (int d = 1; d < kHRGroup; d <<= 1) {
            const int t = __shfl_up(inc, d, kHRGroup);
            if ((lane & (kHRGroup - 1)) >= d) inc += t;
          }
          unsigned m = (w >= w_lo && w <= w_hi) ? full : 0u;
          if (w == w_lo) m &= ~((1u << (c_lo & 31)) - 1u);
          if (c_hi != INT_MAX && w == (c_hi >> 5)) m &= (1u << (c_hi & 31)) - 1u;
          if (m) {
            const int r0 = s_gpre[w / kHRGroup] + inc - pc - (int)s0;
            while (m) {
              const int b = __ffs(m) - 1;
              s_colbuf[r0 + __popc(full & ((1u << b) - 1u))] = w * 32 + b;
              m &= m - 1;
            }
          }
        }
        __syncthreads();
        if (s0 == 0) advance(5);
        for (int t = tid; t < len; t += BLOCK) __builtin_nontemporal_store(s_colbuf[t], &tile.c_col[base + s0 + t]);
        __syncthreads();
        for (int t = tid; t < len; t += BLOCK) s_acc[t] = 0.0;
        __syncthreads();
        if (s0 == 0) advance(6);
        auto add = [&](int j, double v) {
          if (whole || (j >= c_lo && j < c_hi)) atomicAdd(&s_acc[hr_rank(s_bits, s_gpre, j) - (int)s0], v);
        };
        if (one_chunk) {
          RP st = my_st, ln = my_len;
          if (!whole && tid < na) {  // the part of the entry's B row inside the slice's columns
            const RP lo = my_lo;
            const RP hi = c_hi == INT_MAX ? my_len : hr_lower_bound(tile.b_col, my_st, lo, my_len, c_hi);
            st = my_st + lo;
            ln = hi - lo;
            my_lo = hi;
          }
          run_staged(st, ln, add, std::true_type{});
        } else {
          for_each_product<BLOCK, BLOCK, true, T, RP>(tile, a0, a1, tid, s_est, s_epre, s_eval, s_lscan, nullptr, add);
          __syncthreads();
        }
        for (int t = tid; t < len; t += BLOCK) __builtin_nontemporal_store((T)s_acc[t], &tile.c_val[base + s0 + t]);
        __syncthreads();
      }
    } while (false);
    advance(6);  // rows that ended early (symbolic, refused) finish the next row's fetch here
  }
}

// ------------------------------------------------------------------------------------------------------------
// Heavy rows: one workgroup per row, a bitmap of the row's columns (LDS when it fits, else global) and a dense
// float64 accumulator of n entries in HBM (one zeroed slot per resident workgroup; emission re-zeroes it).
// The bitmap gives the sorted column order and the exact count without a sort.
//   MODE 0 symbolic -> nnz[fr]; MODE 1 numeric at off[fr];
//   MODE 2 predicted: write at off[fr] when count <= cap[fr], else bump-allocate; MODE 3 bump-allocate.
// ------------------------------------------------------------------------------------------------------------

// The dense accumulator (and the bitmap when it is in global memory) is private to one workgroup: every access is
// a workgroup-scope atomic (performed in L2) ordered by workgroup barriers, so no device-scope fence is needed.
template <bool LDS_BITMAP>
__device__ inline void heavy_bit_set(unsigned* bits, int j) {
  if (LDS_BITMAP) atomicOr(&bits[j >> 5], 1u << (j & 31));
  else __hip_atomic_fetch_or(&bits[j >> 5], 1u << (j & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <bool LDS_BITMAP>
__device__ inline unsigned heavy_bit_word(unsigned* bits, int64_t w) {
  if (LDS_BITMAP) return bits[w];
  return __hip_atomic_load(&bits[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int MODE, bool LDS_BITMAP, typename T, typename RP>
__global__ __launch_bounds__(kHeavyThreads) void k_heavy(RowsDev<T, RP> R, int bin, double* dense_slots,
                                                         int64_t slot_doubles, unsigned* gbits, int64_t words) {
  extern __shared__ unsigned s_dyn[];
  __shared__ long long s_scan[kHeavyThreads / 64 + 1];
  __shared__ long long s_base;
  constexpr bool NUM = MODE != 0;
  const int tid = threadIdx.x;
  double* dense = dense_slots + (int64_t)blockIdx.x * slot_doubles;
  unsigned* bits = LDS_BITMAP ? s_dyn : gbits + (int64_t)blockIdx.x * words;
  const int count = *R.count_of(bin);
  const int* items = R.bin_items + (int64_t)bin * R.bin_stride;
  for (int idx = blockIdx.x; idx < count; idx += gridDim.x) {
    const int fr = items[idx];
    const TileDev<T, RP> tile = R.tiles[R.blk_tile[fr >> 8]];
    const int64_t i = fr - tile.flat0;
    const int64_t W = ((int64_t)tile.n + 31) / 32;
    if (LDS_BITMAP)
      for (int64_t w = tid; w < W; w += kHeavyThreads) bits[w] = 0u;
    __syncthreads();
    const RP a0 = tile.a_rp[i], a1 = tile.a_rp[i + 1];
    const long long prod = R.prod[fr];
    const long long na = (long long)(a1 - a0);
    const long long avg = (prod + na - 1) / (na > 0 ? na : 1);
    int lg = 0;
    while ((1 << lg) < avg && (1 << lg) < kHeavyThreads) lg++;
    const int L = 1 << lg, S = kHeavyThreads >> lg;
    const int sub = tid >> lg, sl = tid & (L - 1);
    for (RP p = a0 + sub; p < a1; p += S) {
      const int kk = tile.a_col[p];
      const double av = NUM ? (double)tile.a_val[p] : 0.0;
      const RP b0 = tile.b_rp[kk], b1 = tile.b_rp[kk + 1];
      for (RP q = b0 + sl; q < b1; q += L) {
        const int j = tile.b_col[q];
        heavy_bit_set<LDS_BITMAP>(bits, j);
        if (NUM)
          __hip_atomic_fetch_add(&dense[j], av * (double)tile.b_val[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    __syncthreads();
    const int64_t per = (W + kHeavyThreads - 1) / kHeavyThreads;
    const int64_t w0 = tid * per, w1 = (w0 + per < W) ? w0 + per : W;
    long long c = 0;
    for (int64_t w = w0; w < w1; w++) c += __popc(heavy_bit_word<LDS_BITMAP>(bits, w));
    long long total;
    const long long pre = block_excl_scan<kHeavyThreads, long long>(c, &total, s_scan);
    if (MODE == 0) {
      if (tid == 0) R.nnz[fr] = (int)total;
      if (!LDS_BITMAP)
        for (int64_t w = w0; w < w1; w++) __hip_atomic_store(&bits[w], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      if (tid == 0) {
        long long b;
        if (MODE == 1) {
          b = R.off[fr];
        } else if (MODE == 2 && total <= (long long)R.cap[fr]) {
          b = R.off[fr];
        } else {
          b = (long long)atomicAdd(R.cursor, (unsigned long long)total);
          R.off[fr] = b;
        }
        if (b + total > R.out_limit) {  // refused: nothing is written and, after this pass, the row holds 0 entries
          R.stats[3] = 1ull;
          b = -1;
          if (MODE != 1) R.off[fr] = 0;
        }
        s_base = b;
        R.nnz[fr] = (b < 0 && MODE != 1) ? 0 : (int)total;
      }
      __syncthreads();
      const bool fits = s_base >= 0;
      long long pos = s_base + pre;
      for (int64_t w = w0; w < w1; w++) {
        unsigned word = heavy_bit_word<LDS_BITMAP>(bits, w);
        while (word) {  // read and re-zero the accumulator whether or not the row is written
          const int bit = __ffs(word) - 1;
          const int j = (int)(w * 32 + bit);
          const double v = __hip_atomic_exchange(&dense[j], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (fits) {
            __builtin_nontemporal_store(j, &tile.c_col[pos]);
            __builtin_nontemporal_store((T)v, &tile.c_val[pos]);
            pos++;
          }
          word &= word - 1;
        }
        if (!LDS_BITMAP) __hip_atomic_store(&bits[w], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    if (tid == 0) atomicAdd(&R.stats[1], 1ull);
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------------------------
// Scans over flat rows: off[r] = sum_{q < r} cnt[q] (int32 counts -> int64 offsets), off[n] = total.
// Three launches: per-block sums, scan of block sums (one block), per-block scan with base.
// ------------------------------------------------------------------------------------------------------------

constexpr int kScanBlock = 1024;
constexpr int kScanPer = 4;
constexpr int kScanTile = kScanBlock * kScanPer;

// kScanPer consecutive counts of a thread: one 16-byte load inside the array (cnt is 16-byte aligned), else scalars
__device__ inline void scan_load4(const int32_t* cnt, int64_t n, int64_t r0, long long* v) {
  static_assert(kScanPer == 4, "one int4 per thread");
  if (r0 + 3 < n) {
    const int4 x = *reinterpret_cast<const int4*>(cnt + r0);
    v[0] = x.x;
    v[1] = x.y;
    v[2] = x.z;
    v[3] = x.w;
  } else {
#pragma unroll
    for (int q = 0; q < kScanPer; q++) v[q] = r0 + q < n ? cnt[r0 + q] : 0;
  }
}

static __global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int32_t* cnt, int64_t n, long long* partial) {
  __shared__ long long sm[kScanBlock / 64 + 1];
  const int64_t r0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
  long long v[kScanPer];
  scan_load4(cnt, n, r0, v);
  long long s = v[0] + v[1] + v[2] + v[3];
  s = block_sum<kScanBlock, long long>(s, sm);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

static __global__ __launch_bounds__(kScanBlock) void k_scan_top(long long* partial, int64_t nblocks, int64_t* total_out) {
  __shared__ long long sm[kScanBlock / 64 + 1];
  long long run = 0;
  for (int64_t base = 0; base < nblocks; base += kScanBlock) {
    const int64_t r = base + threadIdx.x;
    long long v = r < nblocks ? partial[r] : 0;
    long long tot;
    long long ex = block_excl_scan<kScanBlock, long long>(v, &tot, sm);
    if (r < nblocks) partial[r] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = run;
}

static __global__ __launch_bounds__(kScanBlock) void k_scan_final(const int32_t* cnt, int64_t n, const long long* partial,
                                                           int64_t* off, int64_t off_add) {
  __shared__ long long sm[kScanBlock / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  // each thread owns kScanPer consecutive elements
  const int64_t r0 = base + (int64_t)threadIdx.x * kScanPer;
  long long v[kScanPer];
  scan_load4(cnt, n, r0, v);
  const long long s = v[0] + v[1] + v[2] + v[3];
  long long tot;
  long long ex = block_excl_scan<kScanBlock, long long>(s, &tot, sm);
  long long run = partial[blockIdx.x] + ex + off_add;
  if (r0 + 3 < n) {  // two 16-byte stores (off is 16-byte aligned and r0 a multiple of 4)
    typedef long long ll2 __attribute__((ext_vector_type(2)));
    ll2* o = reinterpret_cast<ll2*>(off + r0);
    o[0] = ll2{run, run + v[0]};
    o[1] = ll2{run + v[0] + v[1], run + v[0] + v[1] + v[2]};
    run += s;
  } else {
    for (int q = 0; q < kScanPer; q++) {
      if (r0 + q < n) off[r0 + q] = run;
      run += v[q];
    }
  }
  if (r0 + kScanPer >= n && r0 < n) off[n] = run;  // the thread holding the last element writes the total
}

// Compaction of the learned path: each row's entries move from the arena (at off[fr], nnz[fr] long) to the final
// CSR (at dst_off[fr]). Work is split by output entries, not rows, so a block of hub rows (tens of thousands of
// entries each) costs the same as any other: block c copies entries [c * kCompactChunk, (c + 1) * kCompactChunk).
constexpr int kCompactChunk = 8192;

// chunk_row[c] = the flat row holding output entry c * kCompactChunk.
static __global__ __launch_bounds__(kRowBlock) void k_chunk_rows(const int64_t* dst_off, int64_t nflat,
                                                                 long long max_out, int32_t* chunk_row) {
  const int64_t r = (int64_t)blockIdx.x * kRowBlock + threadIdx.x;
  if (r >= nflat) return;
  const long long a = dst_off[r], b = dst_off[r + 1] < max_out ? dst_off[r + 1] : max_out;
  for (long long c = (a + kCompactChunk - 1) / kCompactChunk; c * kCompactChunk < b; c++) chunk_row[c] = (int32_t)r;
}

// One block per chunk. The block walks windows of 256 row offsets (LDS) from the chunk's first row; each thread
// finds the row of its entry by binary search in the window and copies it. Loads and stores are unit-stride across
// the wavefront, four entries in flight per thread.
template <typename T, typename RP>
__global__ __launch_bounds__(kRowBlock) void k_compact(RowsDev<T, RP> R, const int64_t* dst_off, int64_t nflat,
                                                       const int32_t* chunk_row, long long max_out, int* dst_col,
                                                       T* dst_val, const int* src_col, const T* src_val) {
  __shared__ long long s_dst[kRowBlock + 1];
  __shared__ long long s_src[kRowBlock];
  const int tid = threadIdx.x;
  long long total = dst_off[nflat];
  if (total > max_out) {  // more entries than the output holds: refuse the excess and flag it
    if (blockIdx.x == 0 && tid == 0) R.stats[3] = 1ull;
    total = max_out;
  }
  const long long c0 = (long long)blockIdx.x * kCompactChunk;
  if (c0 >= total) return;
  const long long c1 = c0 + kCompactChunk < total ? c0 + kCompactChunk : total;
  int64_t r0 = chunk_row[blockIdx.x];
  long long e = c0;
  while (e < c1) {
    const int64_t r = r0 + tid;
    s_dst[tid] = dst_off[r < nflat ? r : nflat];
    s_src[tid] = r < nflat ? R.off[r] : 0;
    if (tid == 0) s_dst[kRowBlock] = dst_off[r0 + kRowBlock < nflat ? r0 + kRowBlock : nflat];
    __syncthreads();
    const long long we = s_dst[kRowBlock] < c1 ? s_dst[kRowBlock] : c1;
    for (long long x0 = e; x0 < we; x0 += 4LL * kRowBlock) {
      long long src[4];
      bool ok[4];
      int a[4];  // largest a with s_dst[a] <= x: four branchless searches in lockstep
#pragma unroll
      for (int q = 0; q < 4; q++) a[q] = 0;
#pragma unroll
      for (int step = kRowBlock / 2; step > 0; step >>= 1) {
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (s_dst[a[q] + step] <= x0 + (long long)q * kRowBlock + tid) a[q] += step;
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const long long x = x0 + (long long)q * kRowBlock + tid;
        ok[q] = x < we;
        src[q] = s_src[a[q]] + (x - s_dst[a[q]]);
      }
      int cv[4];
      T vv[4];
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (ok[q]) {
          // the arena is read once: streaming loads, as the stores
          cv[q] = __builtin_nontemporal_load(&src_col[src[q]]);
          vv[q] = __builtin_nontemporal_load(&src_val[src[q]]);
        }
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (ok[q]) {
          const long long x = x0 + (long long)q * kRowBlock + tid;
          __builtin_nontemporal_store(cv[q], &dst_col[x]);
          __builtin_nontemporal_store(vv[q], &dst_val[x]);
        }
    }
    e = we;
    r0 += kRowBlock;
    __syncthreads();
  }
}

// Per-product row_ptr from flat offsets: rp[i] = off[flat0 + i] - off[flat0] (int32 or int64 out).
template <typename OUT>
__global__ void k_rowptr(const int64_t* off, int64_t flat0, int64_t m, OUT* rp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= m) rp[i] = (OUT)(off[flat0 + i] - off[flat0]);
}

}  // namespace bclx
