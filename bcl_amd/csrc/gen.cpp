// Host-side synthetic matrix generators and tile cutting for the local tile multiply.
//
// Every generator is deterministic in (seed, sizes) and independent of the thread count: each random
// draw comes from a counter-based 64-bit mixer keyed by (seed, stream, counter), so the same matrix is
// produced on any host. The matrices are the workloads the tile multiply is measured on:
//   - uniform:   m x n, exactly `per_row` distinct uniformly random columns per row (Floyd sampling)
//   - banded:    1..max_per_row distinct columns per row inside |i - j| <= bw
//   - R-MAT:     2^scale square, ef * 2^scale edge draws with quadrant probabilities (a, b, c, d),
//                no noise, no vertex permutation, self loops kept, duplicate edges merged
//   - laplacian: 2D 5-point stencil on a g x g grid (4 on the diagonal, -1 for the 4 neighbours)
// Values are U[-1, 1) with 24 random bits, so every value is exact in fp32.
//
// Tile cutting follows CSRMatrix::get_slice_impl_ (bcl/containers/sequential/CSRMatrix.hpp:238-276):
// tile (I, J) of a grid with tile shape (tm, tn) holds rows [I*tm, min((I+1)*tm, m)) and columns
// [J*tn, min((J+1)*tn, n)), re-based to 0, in the source row order. Tile shapes follow
// BlockCustom::seed (bcl/containers/detail/Blocking.hpp:223-232): tm = ceil(m / grid_m).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

inline uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser (Steele, Lea, Flood 2014)
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t counter) {
  return mix64(seed * 0x9E3779B97F4A7C15ULL + stream * 0xD1B54A32D192ED03ULL + counter * 0x8CB92BA72F3D8DD7ULL +
               0x632BE59BD9B4E019ULL);
}

// uniform integer in [0, bound)
inline uint64_t draw_below(uint64_t r, uint64_t bound) {
  return (uint64_t)(((unsigned __int128)r * bound) >> 64);
}

// U[-1, 1) with 24 random bits: exact in fp32
inline float draw_value(uint64_t seed, uint64_t e) {
  uint64_t r = draw(seed, 0x7A1u, e);
  int32_t q = (int32_t)(r >> 40);  // [0, 2^24)
  return (float)(q - (1 << 23)) * (1.0f / (float)(1 << 23));
}

// Floyd's algorithm: `cnt` distinct values in [lo, lo + span), sorted ascending into out.
inline void floyd_sorted(uint64_t seed, uint64_t stream, int64_t lo, int64_t span, int cnt, int32_t* out) {
  int got = 0;
  for (int64_t j = span - cnt; j < span; j++) {
    uint64_t r = draw(seed, stream, (uint64_t)j);
    int64_t t = (int64_t)draw_below(r, (uint64_t)(j + 1));
    bool dup = false;
    for (int q = 0; q < got; q++) {
      if (out[q] == (int32_t)(lo + t)) { dup = true; break; }
    }
    out[got++] = (int32_t)(lo + (dup ? j : t));
  }
  std::sort(out, out + cnt);
}

// Parallel LSD radix sort of 64-bit keys using `bits` significant bits.
void radix_sort_u64(std::vector<uint64_t>& keys, int bits) {
  const int kDigit = 11;
  const int kBuckets = 1 << kDigit;
  size_t n = keys.size();
  std::vector<uint64_t> tmp(n);
  int nthreads = omp_get_max_threads();
  std::vector<size_t> hist((size_t)nthreads * kBuckets);
  for (int shift = 0; shift < bits; shift += kDigit) {
    std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nthreads)
    {
      int t = omp_get_thread_num();
      size_t lo = n * t / nthreads, hi = n * (t + 1) / nthreads;
      size_t* h = &hist[(size_t)t * kBuckets];
      for (size_t i = lo; i < hi; i++) h[(keys[i] >> shift) & (kBuckets - 1)]++;
#pragma omp barrier
#pragma omp single
      {
        size_t run = 0;
        for (int b = 0; b < kBuckets; b++) {
          for (int tt = 0; tt < nthreads; tt++) {
            size_t c = hist[(size_t)tt * kBuckets + b];
            hist[(size_t)tt * kBuckets + b] = run;
            run += c;
          }
        }
      }
      for (size_t i = lo; i < hi; i++) {
        uint64_t k = keys[i];
        tmp[h[(k >> shift) & (kBuckets - 1)]++] = k;
      }
    }
    keys.swap(tmp);
  }
}

}  // namespace

extern "C" {

// m x n, exactly per_row distinct columns per row. row_ptr[m+1] (int64), col[m*per_row], val[m*per_row].
int bclx_gen_uniform(int64_t m, int64_t n, int per_row, uint64_t seed, int64_t* row_ptr, int32_t* col, float* val) {
  if (m < 0 || n <= 0 || per_row < 0 || per_row > n) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++) {
    row_ptr[i] = i * per_row;
    if (per_row) floyd_sorted(seed, (uint64_t)i, 0, n, per_row, col + i * per_row);
    for (int q = 0; q < per_row; q++) val[i * per_row + q] = draw_value(seed ^ 0x55AAu, (uint64_t)(i * per_row + q));
  }
  row_ptr[m] = m * per_row;
  return 0;
}

// Banded: row i has 1..max_per_row distinct columns in [max(0,i-bw), min(n-1,i+bw)].
// Capacity needed for col/val: m * max_per_row. Returns nnz, or -1 on bad arguments.
int64_t bclx_gen_banded(int64_t m, int64_t n, int64_t bw, int max_per_row, uint64_t seed, int64_t* row_ptr,
                        int32_t* col, float* val) {
  if (m < 0 || n <= 0 || bw < 0 || max_per_row <= 0) return -1;
  std::vector<int> cnt(m);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++) {
    int64_t lo = std::max<int64_t>(0, std::min<int64_t>(i, n - 1) - bw);
    int64_t hi = std::min<int64_t>(n - 1, std::min<int64_t>(i, n - 1) + bw);
    int64_t span = hi - lo + 1;
    int c = 1 + (int)draw_below(draw(seed, 0xBA7Du, (uint64_t)i), (uint64_t)max_per_row);
    cnt[i] = (int)std::min<int64_t>(c, span);
  }
  row_ptr[0] = 0;
  for (int64_t i = 0; i < m; i++) row_ptr[i + 1] = row_ptr[i] + cnt[i];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++) {
    int64_t lo = std::max<int64_t>(0, std::min<int64_t>(i, n - 1) - bw);
    int64_t hi = std::min<int64_t>(n - 1, std::min<int64_t>(i, n - 1) + bw);
    floyd_sorted(seed, (uint64_t)i, lo, hi - lo + 1, cnt[i], col + row_ptr[i]);
    for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; e++) val[e] = draw_value(seed ^ 0x55AAu, (uint64_t)e);
  }
  return row_ptr[m];
}

// R-MAT edge list -> deduplicated CSR. n = 2^scale; ef * n edge draws; probabilities (a, b, c, 1-a-b-c).
// Capacity needed for col/val: ef * n. Returns nnz (after merging duplicates), or -1 on bad arguments.
int64_t bclx_gen_rmat(int scale, int ef, double a, double b, double c, uint64_t seed, int64_t* row_ptr,
                      int32_t* col, float* val) {
  if (scale <= 0 || scale > 30 || ef <= 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return -1;
  const int64_t n = (int64_t)1 << scale;
  const int64_t ne = n * (int64_t)ef;
  const uint64_t ta = (uint64_t)(a * 4294967296.0);
  const uint64_t tb = (uint64_t)((a + b) * 4294967296.0);
  const uint64_t tc = (uint64_t)((a + b + c) * 4294967296.0);
  std::vector<uint64_t> keys(ne);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < ne; e++) {
    uint64_t r = 0, cc = 0;
    uint64_t bits = 0;
    for (int lvl = 0; lvl < scale; lvl++) {
      if ((lvl & 1) == 0) bits = draw(seed, 0xED6Eu, (uint64_t)e * 32 + (uint64_t)(lvl >> 1));
      uint64_t u = (lvl & 1) ? (bits >> 32) : (bits & 0xFFFFFFFFULL);
      int q = (u < ta) ? 0 : (u < tb) ? 1 : (u < tc) ? 2 : 3;  // 0:a 1:b (col bit) 2:c (row bit) 3:d
      r = (r << 1) | (uint64_t)(q >> 1);
      cc = (cc << 1) | (uint64_t)(q & 1);
    }
    keys[e] = (r << scale) | cc;
  }
  radix_sort_u64(keys, 2 * scale);
  int64_t nnz = 0;
  for (int64_t e = 0; e < ne; e++)
    if (e == 0 || keys[e] != keys[e - 1]) keys[nnz++] = keys[e];
  std::vector<int64_t> counts(n + 1, 0);
  for (int64_t e = 0; e < nnz; e++) counts[(keys[e] >> scale) + 1]++;
  row_ptr[0] = 0;
  for (int64_t i = 0; i < n; i++) row_ptr[i + 1] = row_ptr[i] + counts[i + 1];
  const uint64_t cmask = (uint64_t)n - 1;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nnz; e++) {
    col[e] = (int32_t)(keys[e] & cmask);
    val[e] = draw_value(seed ^ 0x55AAu, (uint64_t)e);
  }
  return nnz;
}

// 2D 5-point Laplacian on a g x g grid, fp64 values. nnz = 5 g^2 - 4 g. Vertex (x, y) -> x * g + y.
int64_t bclx_gen_laplacian2d(int64_t g, int64_t* row_ptr, int32_t* col, double* val) {
  if (g <= 0 || g * g > INT32_MAX) return -1;
  const int64_t n = g * g;
  row_ptr[0] = 0;
  for (int64_t v = 0; v < n; v++) {
    int64_t x = v / g, y = v % g;
    int deg = (x > 0) + (x < g - 1) + (y > 0) + (y < g - 1);
    row_ptr[v + 1] = row_ptr[v] + 1 + deg;
  }
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < n; v++) {
    int64_t x = v / g, y = v % g;
    int64_t p = row_ptr[v];
    if (x > 0) { col[p] = (int32_t)(v - g); val[p++] = -1.0; }
    if (y > 0) { col[p] = (int32_t)(v - 1); val[p++] = -1.0; }
    col[p] = (int32_t)v; val[p++] = 4.0;
    if (y < g - 1) { col[p] = (int32_t)(v + 1); val[p++] = -1.0; }
    if (x < g - 1) { col[p] = (int32_t)(v + g); val[p++] = -1.0; }
  }
  return row_ptr[n];
}

// Per-tile nnz of a grid cut of a global CSR (int64 row_ptr, sorted int32 columns).
// out_nnz[I * grid_n + J] for tiles of shape (tm, tn). Parallel over tile rows.
void bclx_csr_grid_counts(int64_t m, int64_t n, const int64_t* row_ptr, const int32_t* col, int64_t tm, int64_t tn,
                          int64_t grid_m, int64_t grid_n, int64_t* out_nnz) {
  std::memset(out_nnz, 0, sizeof(int64_t) * grid_m * grid_n);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t I = 0; I < grid_m; I++) {
    int64_t r1 = std::min(m, (I + 1) * tm);
    for (int64_t i = I * tm; i < r1; i++) {
      for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; e++) {
        int64_t J = col[e] / tn;
        if (J < grid_n) out_nnz[I * grid_n + J]++;
      }
    }
  }
}

// Cut tile (I, J): rows [I*tm, min((I+1)*tm, m)), columns [J*tn, min((J+1)*tn, n)).
// t_row_ptr has (rows + 1) int32 entries; t_col/t_val hold the tile's nnz (from bclx_csr_grid_counts).
int bclx_csr_tile_fp32(int64_t m, int64_t n, const int64_t* row_ptr, const int32_t* col, const float* val, int64_t tm,
                       int64_t tn, int64_t I, int64_t J, int32_t* t_row_ptr, int32_t* t_col, float* t_val) {
  int64_t r0 = I * tm, r1 = std::min(m, (I + 1) * tm);
  int64_t c0 = J * tn, c1 = std::min(n, (J + 1) * tn);
  if (r0 >= m || c0 >= n) return -1;
  int64_t p = 0;
  for (int64_t i = r0; i < r1; i++) {
    t_row_ptr[i - r0] = (int32_t)p;
    const int32_t* lo = std::lower_bound(col + row_ptr[i], col + row_ptr[i + 1], (int32_t)c0);
    const int32_t* hi = std::lower_bound(lo, col + row_ptr[i + 1], (int32_t)c1);
    for (const int32_t* q = lo; q < hi; q++) {
      t_col[p] = (int32_t)(*q - c0);
      t_val[p] = val[q - col];
      p++;
    }
  }
  t_row_ptr[r1 - r0] = (int32_t)p;
  return 0;
}

}  // extern "C"
