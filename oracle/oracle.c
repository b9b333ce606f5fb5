/* CPU oracle for the local tile multiply -- TEST INFRASTRUCTURE ONLY (see oracle.h for the contract and the
 * parity pins). Compiled with -ffp-contract=off so every product and sum rounds exactly as written, matching
 * the reference's scalar x86-64 build. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void oracle_csr_free(oracle_csr* c) {
  if (!c) return;
  free(c->row_ptr);
  free(c->col_idx);
  free(c->vals);
  c->row_ptr = NULL;
  c->col_idx = NULL;
  c->vals = NULL;
}

/* ---------------------------------------------------------------------------------------------------------
 * Faithful SpGEMM (bcl/containers/algorithms/spgemm.hpp:40-76 + CSRMatrix::assign CSRMatrix.hpp:106-165).
 * The reference keeps one hash map keyed by the coordinate pair (i, j) with the djb2 hash of the pair's bytes
 * (bcl/core/detail/hash_functions.hpp:10-24). Existing C entries are inserted first (spgemm.hpp:47-54), then
 * for each row i, for each a_ik in CSR order, for each b_kj in CSR order: map[(i,j)] += a_ik * b_kj, where a
 * new entry starts at T{} = +0. Entries are never erased, so exact-zero sums stay. assign() sorts all tuples by
 * (i, j) and builds CSR. The value of an entry is therefore the left-to-right sum, in A-row order of k, of its
 * products, starting from the existing C value or +0.
 * ------------------------------------------------------------------------------------------------------- */

static uint64_t djb2_pair(int32_t i, int32_t j) {
  unsigned char b[8];
  memcpy(b, &i, 4);
  memcpy(b + 4, &j, 4);
  uint64_t h = 5381;
  for (int q = 0; q < 8; q++) h = ((h << 5) + h) + (uint64_t)b[q];
  return h;
}

#define EMPTY_KEY UINT64_MAX

static void radix_sort_keys_perm(uint64_t* keys, int64_t* perm, int64_t n, int bits) {
  uint64_t* k2 = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
  int64_t* p2 = (int64_t*)malloc(sizeof(int64_t) * (n ? n : 1));
  int64_t cnt[65536 + 1];
  for (int shift = 0; shift < bits; shift += 16) {
    memset(cnt, 0, sizeof(cnt));
    for (int64_t q = 0; q < n; q++) cnt[((keys[q] >> shift) & 0xFFFF) + 1]++;
    for (int d = 0; d < 65536; d++) cnt[d + 1] += cnt[d];
    for (int64_t q = 0; q < n; q++) {
      int64_t dst = cnt[(keys[q] >> shift) & 0xFFFF]++;
      k2[dst] = keys[q];
      p2[dst] = perm[q];
    }
    memcpy(keys, k2, sizeof(uint64_t) * n);
    memcpy(perm, p2, sizeof(int64_t) * n);
  }
  free(k2);
  free(p2);
}

static int bits_for(uint64_t x) {
  int b = 0;
  while (b < 64 && (x >> b)) b++;
  return b;
}

#define DEF_FAITHFUL(T, SUF)                                                                                          \
  typedef struct {                                                                                                   \
    uint64_t* keys;                                                                                                  \
    T* vals;                                                                                                         \
    int64_t cap, size;                                                                                               \
  } map_##SUF;                                                                                                       \
  static void map_init_##SUF(map_##SUF* mp, int64_t cap) {                                                           \
    mp->cap = cap;                                                                                                   \
    mp->size = 0;                                                                                                    \
    mp->keys = (uint64_t*)malloc(sizeof(uint64_t) * cap);                                                            \
    mp->vals = (T*)malloc(sizeof(T) * cap);                                                                          \
    for (int64_t q = 0; q < cap; q++) mp->keys[q] = EMPTY_KEY;                                                       \
  }                                                                                                                  \
  static T* map_slot_##SUF(map_##SUF* mp, int32_t i, int32_t j, int insert_value, T v);                              \
  static void map_grow_##SUF(map_##SUF* mp) {                                                                        \
    map_##SUF old = *mp;                                                                                             \
    map_init_##SUF(mp, old.cap * 2);                                                                                 \
    for (int64_t q = 0; q < old.cap; q++)                                                                            \
      if (old.keys[q] != EMPTY_KEY)                                                                                  \
        map_slot_##SUF(mp, (int32_t)(old.keys[q] >> 32), (int32_t)(old.keys[q] & 0xFFFFFFFFu), 1, old.vals[q]);      \
    free(old.keys);                                                                                                  \
    free(old.vals);                                                                                                  \
  }                                                                                                                  \
  /* operator[] (insert_value == 0: new entries start at +0) or insert() (insert_value == 1: keep existing) */       \
  static T* map_slot_##SUF(map_##SUF* mp, int32_t i, int32_t j, int insert_value, T v) {                             \
    if (2 * (mp->size + 1) > mp->cap) map_grow_##SUF(mp);                                                            \
    uint64_t key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;                                                      \
    uint64_t s = djb2_pair(i, j) & (uint64_t)(mp->cap - 1);                                                          \
    for (;;) {                                                                                                       \
      if (mp->keys[s] == key) return &mp->vals[s];                                                                   \
      if (mp->keys[s] == EMPTY_KEY) {                                                                                \
        mp->keys[s] = key;                                                                                           \
        mp->vals[s] = insert_value ? v : (T)0;                                                                       \
        mp->size++;                                                                                                  \
        return &mp->vals[s];                                                                                         \
      }                                                                                                              \
      s = (s + 1) & (uint64_t)(mp->cap - 1);                                                                         \
    }                                                                                                                \
  }                                                                                                                  \
  static int faithful_##SUF(int64_t m, int64_t n, const int64_t* a_rp, const int32_t* a_col, const void* a_val_,     \
                            int a_is_f32, const int64_t* b_rp, const int32_t* b_col, const void* b_val_,             \
                            const oracle_csr* c0, int dead_row_alloc, oracle_csr* out) {                             \
    map_##SUF mp;                                                                                                    \
    map_init_##SUF(&mp, 1024);                                                                                       \
    if (c0 && c0->nnz > 0) {                                                                                         \
      const T* cv = (const T*)c0->vals;                                                                              \
      for (int64_t i = 0; i < c0->m; i++)                                                                            \
        for (int64_t p = c0->row_ptr[i]; p < c0->row_ptr[i + 1]; p++)                                                \
          map_slot_##SUF(&mp, (int32_t)i, c0->col_idx[p], 1, cv[p]);                                                 \
    }                                                                                                                \
    for (int64_t i = 0; i < m; i++) {                                                                                \
      volatile uint64_t* dead = NULL;                                                                                \
      if (dead_row_alloc) {                                                                                          \
        dead = (volatile uint64_t*)calloc((size_t)((n + 63) / 64) + 1, sizeof(uint64_t));                            \
        dead[0] = 0;                                                                                                 \
      }                                                                                                              \
      for (int64_t kp = a_rp[i]; kp < a_rp[i + 1]; kp++) {                                                           \
        int32_t kk = a_col[kp];                                                                                      \
        T av = a_is_f32 ? (T)((const float*)a_val_)[kp] : (T)((const double*)a_val_)[kp];                            \
        for (int64_t jp = b_rp[kk]; jp < b_rp[kk + 1]; jp++) {                                                       \
          T bv = a_is_f32 ? (T)((const float*)b_val_)[jp] : (T)((const double*)b_val_)[jp];                          \
          T prod = av * bv;                                                                                          \
          T* slot = map_slot_##SUF(&mp, (int32_t)i, b_col[jp], 0, (T)0);                                             \
          *slot = *slot + prod;                                                                                      \
        }                                                                                                            \
      }                                                                                                              \
      if (dead) free((void*)dead);                                                                                   \
    }                                                                                                                \
    int64_t nnz = mp.size;                                                                                           \
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));                                          \
    int64_t* perm = (int64_t*)malloc(sizeof(int64_t) * (nnz ? nnz : 1));                                             \
    int64_t q = 0;                                                                                                   \
    for (int64_t s = 0; s < mp.cap; s++)                                                                             \
      if (mp.keys[s] != EMPTY_KEY) {                                                                                 \
        keys[q] = mp.keys[s];                                                                                        \
        perm[q] = s;                                                                                                 \
        q++;                                                                                                         \
      }                                                                                                              \
    radix_sort_keys_perm(keys, perm, nnz, 32 + bits_for((uint64_t)(m > 0 ? m : 1)));                                 \
    out->m = m;                                                                                                      \
    out->n = n;                                                                                                      \
    out->nnz = nnz;                                                                                                  \
    out->row_ptr = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));                                                 \
    out->col_idx = (int32_t*)malloc(sizeof(int32_t) * (nnz ? nnz : 1));                                              \
    out->vals = malloc(sizeof(T) * (nnz ? nnz : 1));                                                                 \
    T* ov = (T*)out->vals;                                                                                           \
    int64_t r = 0;                                                                                                   \
    for (int64_t c = 0; c < nnz; c++) {                                                                              \
      int64_t i = (int64_t)(keys[c] >> 32);                                                                          \
      out->col_idx[c] = (int32_t)(keys[c] & 0xFFFFFFFFu);                                                            \
      ov[c] = mp.vals[perm[c]];                                                                                      \
      while (r < i) out->row_ptr[++r] = c;                                                                           \
    }                                                                                                                \
    for (; r < m; r++) out->row_ptr[r + 1] = nnz;                                                                    \
    free(keys);                                                                                                      \
    free(perm);                                                                                                      \
    free(mp.keys);                                                                                                   \
    free(mp.vals);                                                                                                   \
    return 0;                                                                                                        \
  }

DEF_FAITHFUL(float, f32)
DEF_FAITHFUL(double, f64)

int oracle_spgemm_f32(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col, const float* a_val,
                      const int64_t* b_rp, const int32_t* b_col, const float* b_val, const oracle_csr* c0,
                      int dead_row_alloc, oracle_csr* out) {
  (void)k;
  return faithful_f32(m, n, a_rp, a_col, a_val, 1, b_rp, b_col, b_val, c0, dead_row_alloc, out);
}

int oracle_spgemm_f64(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col, const double* a_val,
                      const int64_t* b_rp, const int32_t* b_col, const double* b_val, const oracle_csr* c0,
                      int dead_row_alloc, oracle_csr* out) {
  (void)k;
  return faithful_f64(m, n, a_rp, a_col, a_val, 0, b_rp, b_col, b_val, c0, dead_row_alloc, out);
}

int oracle_spgemm_f32_to_f64(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col,
                             const float* a_val, const int64_t* b_rp, const int32_t* b_col, const float* b_val,
                             oracle_csr* out) {
  (void)k;
  return faithful_f64(m, n, a_rp, a_col, a_val, 1, b_rp, b_col, b_val, NULL, 0, out);
}

/* ---------------------------------------------------------------------------------------------------------
 * Row-parallel hash SpGEMM: SimpleHash (bcl/containers/sequential/SimpleHash.hpp:24-165) per row, reserved to
 * pow2 >= max(16, products_i) (SimpleHash::reserve :133-146), probe (col * 107 + p) & (cap - 1)
 * (hash_functions.hpp:27-31, SimpleHash.hpp:88-119, :154-156), then sort the row by column
 * (SparseHashAccumulator.hpp:55-63). Two passes: count, then fill.
 * ------------------------------------------------------------------------------------------------------- */

static int64_t pow2_at_least(int64_t x) {
  int64_t c = 16;
  while (c < x) c <<= 1;
  return c;
}

typedef struct {
  int32_t col;
  int32_t pad;
  double v;
} colval;

static int cmp_colval(const void* a, const void* b) {
  int32_t x = ((const colval*)a)->col, y = ((const colval*)b)->col;
  return (x > y) - (x < y);
}

static void sort_colval(colval* e, int64_t cnt) {
  if (cnt <= 32) {
    for (int64_t s = 1; s < cnt; s++) {
      colval t = e[s];
      int64_t q = s - 1;
      while (q >= 0 && e[q].col > t.col) {
        e[q + 1] = e[q];
        q--;
      }
      e[q + 1] = t;
    }
  } else {
    qsort(e, (size_t)cnt, sizeof(colval), cmp_colval);
  }
}

int oracle_spgemm_row_nnz(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col,
                          const int64_t* b_rp, const int32_t* b_col, int nthreads, int64_t* out, int64_t* prod) {
  (void)k;
  (void)n;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    int64_t cap = 0;
    int32_t* keys = NULL;
    unsigned char* occ = NULL;
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < m; i++) {
      int64_t pr = 0;
      for (int64_t kp = a_rp[i]; kp < a_rp[i + 1]; kp++) pr += b_rp[a_col[kp] + 1] - b_rp[a_col[kp]];
      if (prod) prod[i] = pr;
      int64_t want = pow2_at_least(pr);
      if (want > cap) {
        free(keys);
        free(occ);
        cap = want;
        keys = (int32_t*)malloc(sizeof(int32_t) * cap);
        occ = (unsigned char*)malloc((size_t)cap);
      }
      memset(occ, 0, (size_t)want);
      int64_t cnt = 0;
      for (int64_t kp = a_rp[i]; kp < a_rp[i + 1]; kp++) {
        int32_t kk = a_col[kp];
        for (int64_t jp = b_rp[kk]; jp < b_rp[kk + 1]; jp++) {
          int32_t j = b_col[jp];
          uint64_t s = ((uint64_t)(uint32_t)j * 107u) & (uint64_t)(want - 1);
          while (occ[s] && keys[s] != j) s = (s + 1) & (uint64_t)(want - 1);
          if (!occ[s]) {
            occ[s] = 1;
            keys[s] = j;
            cnt++;
          }
        }
      }
      out[i] = cnt;
    }
    free(keys);
    free(occ);
  }
  return 0;
}

#define DEF_ROWHASH(T, SUF)                                                                                           \
  int oracle_spgemm_rowhash_##SUF(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col,        \
                                  const T* a_val, const int64_t* b_rp, const int32_t* b_col, const T* b_val,         \
                                  int nthreads, oracle_csr* out) {                                                   \
    int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * (m ? m : 1));                                                  \
    oracle_spgemm_row_nnz(m, k, n, a_rp, a_col, b_rp, b_col, nthreads, cnt, NULL);                                   \
    out->m = m;                                                                                                      \
    out->n = n;                                                                                                      \
    out->row_ptr = (int64_t*)malloc(sizeof(int64_t) * (m + 1));                                                      \
    out->row_ptr[0] = 0;                                                                                             \
    for (int64_t i = 0; i < m; i++) out->row_ptr[i + 1] = out->row_ptr[i] + cnt[i];                                  \
    free(cnt);                                                                                                       \
    out->nnz = out->row_ptr[m];                                                                                      \
    out->col_idx = (int32_t*)malloc(sizeof(int32_t) * (out->nnz ? out->nnz : 1));                                    \
    out->vals = malloc(sizeof(T) * (out->nnz ? out->nnz : 1));                                                       \
    T* ov = (T*)out->vals;                                                                                           \
    _Pragma("omp parallel num_threads(nthreads > 0 ? nthreads : 1)") {                                              \
      int64_t cap = 0;                                                                                               \
      colval* tab = NULL;                                                                                            \
      unsigned char* occ = NULL;                                                                                     \
      _Pragma("omp for schedule(dynamic, 64)") for (int64_t i = 0; i < m; i++) {                                     \
        int64_t pr = 0;                                                                                              \
        for (int64_t kp = a_rp[i]; kp < a_rp[i + 1]; kp++) pr += b_rp[a_col[kp] + 1] - b_rp[a_col[kp]];              \
        int64_t want = pow2_at_least(pr);                                                                            \
        if (want > cap) {                                                                                            \
          free(tab);                                                                                                 \
          free(occ);                                                                                                 \
          cap = want;                                                                                                \
          tab = (colval*)malloc(sizeof(colval) * cap);                                                               \
          occ = (unsigned char*)malloc((size_t)cap);                                                                 \
        }                                                                                                            \
        memset(occ, 0, (size_t)want);                                                                                \
        for (int64_t kp = a_rp[i]; kp < a_rp[i + 1]; kp++) {                                                         \
          int32_t kk = a_col[kp];                                                                                    \
          T av = a_val[kp];                                                                                          \
          for (int64_t jp = b_rp[kk]; jp < b_rp[kk + 1]; jp++) {                                                     \
            int32_t j = b_col[jp];                                                                                   \
            T p = av * b_val[jp];                                                                                    \
            uint64_t s = ((uint64_t)(uint32_t)j * 107u) & (uint64_t)(want - 1);                                      \
            while (occ[s] && tab[s].col != j) s = (s + 1) & (uint64_t)(want - 1);                                    \
            if (!occ[s]) {                                                                                           \
              occ[s] = 1;                                                                                            \
              tab[s].col = j;                                                                                        \
              tab[s].v = (double)((T)0 + p);                                                                         \
            } else {                                                                                                 \
              tab[s].v = (double)((T)tab[s].v + p);                                                                  \
            }                                                                                                        \
          }                                                                                                          \
        }                                                                                                            \
        int64_t q = 0;                                                                                               \
        for (int64_t s = 0; s < want; s++)                                                                           \
          if (occ[s]) tab[q++] = tab[s];                                                                             \
        sort_colval(tab, q);                                                                                         \
        int64_t base = out->row_ptr[i];                                                                              \
        for (int64_t e = 0; e < q; e++) {                                                                            \
          out->col_idx[base + e] = tab[e].col;                                                                       \
          ov[base + e] = (T)tab[e].v;                                                                                \
        }                                                                                                            \
      }                                                                                                              \
      free(tab);                                                                                                     \
      free(occ);                                                                                                     \
    }                                                                                                                \
    return 0;                                                                                                        \
  }

DEF_ROWHASH(float, f32)
DEF_ROWHASH(double, f64)

/* ---------------------------------------------------------------------------------------------------------
 * SpMM (spmm_wrapper_generic, bcl/containers/algorithms/spgemm.hpp:78-96, called with plus/multiplies by
 * spmm_wrapper :140-149): for i, for k_ptr in A row i, for j < n: c[i*ldc + j] = c + v * b[k*ldb + j].
 * ------------------------------------------------------------------------------------------------------- */
void oracle_spmm_f32(int64_t m, int64_t n, int64_t k, const int64_t* a_rp, const int32_t* a_col, const float* a_val,
                     const float* b, int64_t ldb, float* c, int64_t ldc) {
  (void)k;
  for (int64_t i = 0; i < m; i++)
    for (int64_t kp = a_rp[i]; kp < a_rp[i + 1]; kp++) {
      float v = a_val[kp];
      int64_t kk = a_col[kp];
      for (int64_t j = 0; j < n; j++) c[i * ldc + j] = c[i * ldc + j] + v * b[kk * ldb + j];
    }
}

static double bf16_to_double(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return (double)f;
}

void oracle_spmm_f64(int64_t m, int64_t n, int64_t k, const int64_t* a_rp, const int32_t* a_col, const float* a_val,
                     const float* b, const uint16_t* b_bf16, int64_t ldb, const float* c0, int64_t ldc, double* c64,
                     int nthreads) {
  (void)k;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < m; i++) {
    double* crow = c64 + i * ldc;
    for (int64_t j = 0; j < n; j++) crow[j] = c0 ? (double)c0[i * ldc + j] : 0.0;
    for (int64_t kp = a_rp[i]; kp < a_rp[i + 1]; kp++) {
      double v = (double)a_val[kp];
      int64_t kk = a_col[kp];
      if (b_bf16) {
        for (int64_t j = 0; j < n; j++) crow[j] += v * bf16_to_double(b_bf16[kk * ldb + j]);
      } else {
        for (int64_t j = 0; j < n; j++) crow[j] += v * (double)b[kk * ldb + j];
      }
    }
  }
}

/* Dense C += A * B in float64 for rows [r0, r1), row-major (cblas_{s,d}gemm RowMajor NoTrans, alpha = beta = 1,
 * as called by gemm_cowns, bcl/containers/algorithms/gemm.hpp:86-107 and cblas_wrapper.hpp:16-40). */
void oracle_gemm_f64(int64_t m, int64_t n, int64_t k, const float* a, int64_t lda, const float* b, int64_t ldb,
                     double* c, int64_t ldc, int64_t r0, int64_t r1, int nthreads) {
  (void)m;
  const int64_t RB = 8;
  int64_t nblk = (r1 - r0 + RB - 1) / RB;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t blk = 0; blk < nblk; blk++) {
    int64_t i0 = r0 + blk * RB, i1 = i0 + RB < r1 ? i0 + RB : r1;
    for (int64_t p = 0; p < k; p++) {
      const float* brow = b + p * ldb;
      for (int64_t i = i0; i < i1; i++) {
        double av = (double)a[i * lda + p];
        double* crow = c + i * ldc;
        for (int64_t j = 0; j < n; j++) crow[j] += av * (double)brow[j];
      }
    }
  }
}

/* SparseHashAccumulator::get_matrix (SparseHashAccumulator.hpp:101-182) for tiles at offset (0, 0): per row a
 * SimpleHash reserved to pow2 >= max(16, sum of the row's nnz over tiles); values_[idx] = values_[idx] + v in
 * tile order (a new entry starts at T{} = +0); the row is sorted by column; CSR by scan. */
int oracle_accumulate_f32(int64_t m, int64_t n, int ntiles, const int64_t* const* rps, const int32_t* const* cols,
                          const float* const* vals, int nthreads, oracle_csr* out) {
  int64_t* cnt = (int64_t*)calloc((size_t)(m ? m : 1), sizeof(int64_t));
  colval** rows = (colval**)calloc((size_t)(m ? m : 1), sizeof(colval*));
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    int64_t cap = 0;
    colval* tab = NULL;
    unsigned char* occ = NULL;
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < m; i++) {
      int64_t est = 0;
      for (int t = 0; t < ntiles; t++) est += rps[t][i + 1] - rps[t][i];
      int64_t want = pow2_at_least(est);
      if (want > cap) {
        free(tab);
        free(occ);
        cap = want;
        tab = (colval*)malloc(sizeof(colval) * cap);
        occ = (unsigned char*)malloc((size_t)cap);
      }
      memset(occ, 0, (size_t)want);
      for (int t = 0; t < ntiles; t++) {
        for (int64_t p = rps[t][i]; p < rps[t][i + 1]; p++) {
          int32_t j = cols[t][p];
          if (j < 0 || j >= n) continue;
          uint64_t s = ((uint64_t)(uint32_t)j * 107u) & (uint64_t)(want - 1);
          while (occ[s] && tab[s].col != j) s = (s + 1) & (uint64_t)(want - 1);
          if (!occ[s]) {
            occ[s] = 1;
            tab[s].col = j;
            tab[s].v = (double)(0.0f + vals[t][p]);
          } else {
            tab[s].v = (double)((float)tab[s].v + vals[t][p]);
          }
        }
      }
      int64_t q = 0;
      for (int64_t s = 0; s < want; s++)
        if (occ[s]) q++;
      rows[i] = (colval*)malloc(sizeof(colval) * (q ? q : 1));
      q = 0;
      for (int64_t s = 0; s < want; s++)
        if (occ[s]) rows[i][q++] = tab[s];
      sort_colval(rows[i], q);
      cnt[i] = q;
    }
    free(tab);
    free(occ);
  }
  out->m = m;
  out->n = n;
  out->row_ptr = (int64_t*)malloc(sizeof(int64_t) * (m + 1));
  out->row_ptr[0] = 0;
  for (int64_t i = 0; i < m; i++) out->row_ptr[i + 1] = out->row_ptr[i] + cnt[i];
  out->nnz = out->row_ptr[m];
  out->col_idx = (int32_t*)malloc(sizeof(int32_t) * (out->nnz ? out->nnz : 1));
  out->vals = malloc(sizeof(float) * (out->nnz ? out->nnz : 1));
  float* ov = (float*)out->vals;
  for (int64_t i = 0; i < m; i++) {
    for (int64_t e = 0; e < cnt[i]; e++) {
      out->col_idx[out->row_ptr[i] + e] = rows[i][e].col;
      ov[out->row_ptr[i] + e] = (float)rows[i][e].v;
    }
    free(rows[i]);
  }
  free(rows);
  free(cnt);
  return 0;
}

/* MLP 16 -> 64 -> 64 -> 1 in float64. */
void oracle_mlp_forward_backward(int64_t nrows, const float* x, const float* target, const float* params, double* y,
                                 double* grads, double* loss) {
  const float* W1 = params;
  const float* b1 = W1 + 64 * 16;
  const float* W2 = b1 + 64;
  const float* b2 = W2 + 64 * 64;
  const float* W3 = b2 + 64;
  const float* b3 = W3 + 64;
  double l = 0.0;
  if (grads) memset(grads, 0, sizeof(double) * 5313);
  for (int64_t r = 0; r < nrows; r++) {
    const float* xr = x + r * 16;
    double z1[64], h1[64], z2[64], h2[64];
    for (int o = 0; o < 64; o++) {
      double s = (double)b1[o];
      for (int q = 0; q < 16; q++) s += (double)W1[o * 16 + q] * (double)xr[q];
      z1[o] = s;
      h1[o] = s > 0 ? s : 0;
    }
    for (int o = 0; o < 64; o++) {
      double s = (double)b2[o];
      for (int q = 0; q < 64; q++) s += (double)W2[o * 64 + q] * h1[q];
      z2[o] = s;
      h2[o] = s > 0 ? s : 0;
    }
    double out = (double)b3[0];
    for (int q = 0; q < 64; q++) out += (double)W3[q] * h2[q];
    if (y) y[r] = out;
    if (target) {
      double d = out - (double)target[r];
      l += 0.5 * d * d;
      if (grads) {
        double g = d / (double)nrows;
        double* gW1 = grads;
        double* gb1 = gW1 + 64 * 16;
        double* gW2 = gb1 + 64;
        double* gb2 = gW2 + 64 * 64;
        double* gW3 = gb2 + 64;
        double* gb3 = gW3 + 64;
        double d2[64], d1[64];
        gb3[0] += g;
        for (int q = 0; q < 64; q++) {
          gW3[q] += g * h2[q];
          d2[q] = z2[q] > 0 ? g * (double)W3[q] : 0.0;
        }
        for (int o = 0; o < 64; o++) {
          gb2[o] += d2[o];
          for (int q = 0; q < 64; q++) gW2[o * 64 + q] += d2[o] * h1[q];
        }
        for (int q = 0; q < 64; q++) {
          double s = 0;
          for (int o = 0; o < 64; o++) s += d2[o] * (double)W2[o * 64 + q];
          d1[q] = z1[q] > 0 ? s : 0.0;
        }
        for (int o = 0; o < 64; o++) {
          gb1[o] += d1[o];
          for (int q = 0; q < 16; q++) gW1[o * 16 + q] += d1[o] * (double)xr[q];
        }
      }
    }
  }
  if (loss) *loss = nrows ? l / (double)nrows : 0.0;
}
