/* bclx -- MI355X-native local tile multiply for BCL's distributed gemm (C ABI).
 *
 * Every entry point takes plain pointers and sizes, returns a bclx_status and never throws. Pointers named
 * d_* are device (HBM) pointers, h_* are host pointers. Calls are ordered on the handle's HIP stream;
 * functions that report a size to the host (nnz of a product) synchronise that stream. One handle per host
 * thread.
 *
 * Reference interfaces replaced (paths relative to the BCL source tree):
 *   bclx_spgemm_symbolic + bclx_spgemm_numeric
 *       BCL::gemm(const CSRMatrix&, const CSRMatrix&, CSRMatrix&)  bcl/containers/algorithms/spgemm.hpp:40-76
 *       as reached through CSRMatrix::dot                          bcl/containers/sequential/CSRMatrix.hpp:229-235
 *       (GPU counterpart: spgemm_cusparse                          bcl/containers/experimental/cuda/sequential/
 *                                                                  cusparse_util.cuh:269-514)
 *   bclx_accumulate_tiles
 *       SparseHashAccumulator::get_matrix                          bcl/containers/sequential/SparseHashAccumulator.hpp:101-182
 *       also the C += part of BCL::gemm (pre-insert of C, spgemm.hpp:47-54: C + A*B is the sum of two tiles)
 *       and the pairwise sum_cusparse (cusparse_util.cuh:86-249)
 *   bclx_stage_*  (batched SUMMA stage)
 *       the per-stage local products of BCL::gemm(SPMatrix, SPMatrix, SPMatrix)  spgemm.hpp:181-242
 *   bclx_spmm
 *       spmm_wrapper(m, n, k, nnz, c, b, a_values, a_rowptr, a_colind, ldb, ldc)  spgemm.hpp:140-149
 *       (GPU counterpart: spmm_cusparse, cusparse_util.cuh:675-775)
 *   bclx_gemm
 *       cblas_gemm_wrapper_(layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
 *                                                                  bcl/containers/algorithms/cblas_wrapper.hpp:16-40
 */
#ifndef BCLX_H
#define BCLX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BCLX_API __attribute__((visibility("default")))

typedef enum {
  BCLX_OK = 0,
  BCLX_ERR_INVALID = 1,     /* bad argument (shape mismatch, null pointer, unsupported width) */
  BCLX_ERR_HIP = 2,         /* a HIP runtime call failed */
  BCLX_ERR_ALLOC = 3,       /* device allocation failed */
  BCLX_ERR_OVERFLOW = 4,    /* a result does not fit the requested index width or a fixed capacity */
  BCLX_ERR_UNSUPPORTED = 5  /* valid request this build does not implement */
} bclx_status;

typedef enum { BCLX_F32 = 0, BCLX_F64 = 1, BCLX_BF16 = 2 } bclx_dtype;

/* 0-based CSR matrix in device memory. row_ptr has m + 1 entries of ptr_width bytes (4 or 8) starting at 0;
 * col_idx holds int32 columns, sorted and unique within each row; vals holds nnz values of dtype. */
typedef struct {
  int64_t m, n, nnz;
  int32_t ptr_width;
  int32_t dtype;
  void* row_ptr;
  int32_t* col_idx;
  void* vals;
} bclx_csr;

typedef struct bclx_handle bclx_handle;
typedef struct bclx_spgemm_plan bclx_spgemm_plan;
typedef struct bclx_stage_plan bclx_stage_plan;

#define BCLX_TUNING_LEN 6

/* Last error message of the calling thread (empty string when none). */
BCLX_API const char* bclx_last_error(void);

/* Handle on `device` ordered on `stream` (a hipStream_t; NULL = the default stream). */
BCLX_API bclx_status bclx_create(int device, void* stream, bclx_handle** out);
BCLX_API bclx_status bclx_destroy(bclx_handle* h);
BCLX_API bclx_status bclx_set_stream(bclx_handle* h, void* stream);
BCLX_API bclx_status bclx_synchronize(bclx_handle* h);

/* Device allocator of a handle, for the library's own buffers (plans, scratch) and for the outputs bclx_spgemm
 * allocates. alloc(bytes, user) returns device memory or NULL (the call that needed it then fails with
 * BCLX_ERR_ALLOC); free(ptr, user) releases it. Both NULL: hipMalloc / hipFree (the default). Replaces BCL's
 * Allocator template parameter (bcl/containers/sequential/CSRMatrix.hpp:32-36, rebound for the index arrays at
 * :41-42; bcl_allocator, bcl/core/alloc.hpp:60-105). Buffers are freed through the allocator that made them, so its
 * functions must stay valid until every plan made under them is destroyed. */
typedef void* (*bclx_alloc_fn)(size_t bytes, void* user);
typedef void (*bclx_free_fn)(void* ptr, void* user);
BCLX_API bclx_status bclx_set_allocator(bclx_handle* h, bclx_alloc_fn alloc, bclx_free_fn free_fn, void* user);

/* ---- SpGEMM C = A * B, two phases (exact-zero sums are kept, as BCL does) ---------------------------------
 * symbolic: row analysis (products per row), binning and an exact per-row count of C's pattern; returns
 * nnz(C). numeric: fills caller-allocated C (row_ptr m+1 of C->ptr_width, col_idx/vals of nnz) with sorted
 * columns and values accumulated in float64 and rounded once to C->dtype. A and B must share dtype and
 * ptr_width. C->ptr_width 4 fails with BCLX_ERR_OVERFLOW when nnz(C) > 2^31 - 1. */
BCLX_API bclx_status bclx_spgemm_symbolic(bclx_handle* h, const bclx_csr* A, const bclx_csr* B,
                                          bclx_spgemm_plan** plan, int64_t* nnz_c);
BCLX_API bclx_status bclx_spgemm_numeric(bclx_spgemm_plan* plan, bclx_csr* C);
BCLX_API bclx_status bclx_spgemm_plan_destroy(bclx_spgemm_plan* plan);
/* C = A * B in one call with library-allocated outputs (the way CSRMatrix::dot returns a fresh matrix,
 * bcl/containers/sequential/CSRMatrix.hpp:229-235): C->row_ptr (m+1 entries of C->ptr_width), C->col_idx and C->vals
 * (nnz entries each, at least one element) are allocated with the handle's allocator and owned by the caller, who
 * releases them with the matching free function (hipFree by default). Sets C->m, C->n, C->nnz and C->dtype. */
BCLX_API bclx_status bclx_spgemm(bclx_handle* h, const bclx_csr* A, const bclx_csr* B, bclx_csr* C);
/* C = A * B in one call that allocates no device memory, for callers that cannot call hipMalloc (code under HIP
 * graph capture, serving loops with fixed memory, harnesses that forbid allocation). No counterpart in the reference.
 * C->row_ptr (A->m + 1 entries of C->ptr_width), C->col_idx and C->vals are the caller's; on entry C->nnz is how many
 * entries they hold, on return nnz(C). A result that does not fit fails with BCLX_ERR_OVERFLOW and C->nnz set to the
 * entries needed. The per-row workspace lives in a fixed 64 MiB device array that each GPU holds until the process
 * ends: A may have about 600 K rows at most (more fails with BCLX_ERR_OVERFLOW), and B at most 2^19 columns (more
 * fails with BCLX_ERR_UNSUPPORTED). Calls that overlap on one GPU share that array and race. Synchronises the
 * handle's stream. */
BCLX_API bclx_status bclx_spgemm_noalloc(bclx_handle* h, const bclx_csr* A, const bclx_csr* B, bclx_csr* C);
/* Per-row products (int64, m entries) and, after symbolic, nnz(C_i) (int32) of a plan, copied to host. */
BCLX_API bclx_status bclx_spgemm_plan_rows(bclx_spgemm_plan* plan, int64_t* h_products, int32_t* h_row_nnz);

/* ---- sparse add / cross-stage accumulation --------------------------------------------------------------
 * C = sum of ntiles m x n CSR matrices, summed per entry in tile order in float64; columns sorted.
 * Two calls: with C->col_idx == NULL only C->nnz (and C->row_ptr when non-NULL) are produced. */
BCLX_API bclx_status bclx_accumulate_tiles(bclx_handle* h, int ntiles, const bclx_csr* tiles, bclx_csr* C);

/* ---- batched SUMMA stage ----------------------------------------------------------------------------------
 * npairs products C_p = A_p * B_p run as one stage. mode 0 = plain (per-product two-phase launches, host
 * synchronised like a per-tile library call); mode 1 = batched + learned: one persistent pipeline over all
 * products, per-row output sizes predicted by the row-work MLP in place of the symbolic pass, overflow rows
 * redone exactly, captured in a HIP graph and replayed by bclx_stage_execute. Outputs live in the plan. */
typedef struct {
  int32_t mode;             /* 0 plain, 1 batched + learned, 2 batched + exact symbolic (no predictor) */
  int32_t use_graph;        /* mode 1/2: capture the pipeline in a HIP graph (1) or launch eagerly (0) */
  float margin;             /* predicted capacity = ceil(pred * (1 + margin)) + slack */
  int32_t slack;
  const float* h_mlp_params; /* 5313 fp32 MLP parameters (NULL: mode 1 falls back to upper bounds) */
  int32_t persistent_blocks; /* reserved: pass 0 (grids are sized from the plan's own bin counts) */
  const int32_t* h_tuning;   /* optional BCLX_TUNING_LEN binning knobs (autotuned), NULL = defaults:
                              *   [0] work_per_lane (64)  lanes per row >= products / work_per_lane
                              *   [1] heavy_need (4096)   rows needing more slots use the heavy accumulator
                              *   [2] load_shift (1)      hash slots >= need << load_shift
                              *   [3] esc_min (32)        rows with esc_min < products <= esc_max are sorted
                              *   [4] esc_max (4096)      instead of hashed (0 turns sorting off; max 4096)
                              *   [5] bin_streams (1)     concurrent graph branches for the bins (1..8) */
} bclx_stage_opts;

BCLX_API bclx_status bclx_stage_plan_create(bclx_handle* h, int npairs, const bclx_csr* A, const bclx_csr* B,
                                            const bclx_stage_opts* opts, bclx_stage_plan** plan);
BCLX_API bclx_status bclx_stage_execute(bclx_stage_plan* plan);
/* Device view of product p's result (row_ptr int64, pointers into the plan's arena; valid until destroy).
 * A plan is sized for the sparsity structure of its inputs; values may change between executions. If the structure
 * changes, an execution that still fits the plan's bins and capacities recomputes the new products; one that does
 * not refuses the rows it cannot place, and this call then fails with BCLX_ERR_OVERFLOW (make a new plan). Mode 0
 * fails in bclx_stage_execute instead. */
BCLX_API bclx_status bclx_stage_result(bclx_stage_plan* plan, int p, bclx_csr* C);
/* Stats: [0] total products, [1] total nnz(C), [2] overflow rows, [3] predicted rows, [4] arena capacity,
 * [5] rows, [6] heavy rows, [7] sorted (ESC) rows. Fails with BCLX_ERR_OVERFLOW, as bclx_stage_result does, when the
 * last execution refused rows for lack of room. */
BCLX_API bclx_status bclx_stage_stats(bclx_stage_plan* plan, int64_t* h_stats, int nstats);
/* The row predictions of the last execution of a learned plan (mode 1, made with network weights): for each of the
 * first max rows the network sized, the product p, the row within it, the network's predicted entry count, the
 * capacity that row was given and its exact entry count (any array may be NULL); *count = rows the network sized.
 * Diagnostic: it reads back what the graph computed. */
BCLX_API bclx_status bclx_stage_predictions(bclx_stage_plan* plan, int64_t max, int32_t* h_product, int32_t* h_row,
                                            float* h_pred, int32_t* h_cap, int32_t* h_nnz, int64_t* count);
BCLX_API bclx_status bclx_stage_plan_destroy(bclx_stage_plan* plan);

/* ---- learned row-work predictor (MLP 16 -> 64 -> 64 -> 1, ReLU) -------------------------------------------
 * Parameter layout: W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3[1] = 5313 floats. */
#define BCLX_MLP_NPARAM 5313
#define BCLX_MLP_NFEAT 16
/* Features of every row of C = A * B (d_feat: m x 16 fp32), plus products per row (d_prod int64, m). */
BCLX_API bclx_status bclx_row_features(bclx_handle* h, const bclx_csr* A, const bclx_csr* B, float* d_feat,
                                       int64_t* d_prod);
BCLX_API bclx_status bclx_mlp_forward(bclx_handle* h, int64_t nrows, const float* d_x, const float* d_params,
                                      float* d_y);
/* One minibatch: d_grads (5313 fp32) = gradient of 0.5 * mean((y - t)^2); *d_loss = that loss (fp32). */
BCLX_API bclx_status bclx_mlp_grad(bclx_handle* h, int64_t nrows, const float* d_x, const float* d_t,
                                   const float* d_params, float* d_grads, float* d_loss);
/* Adam step on d_params with moments d_m, d_v (5313 each), step number t >= 1. */
BCLX_API bclx_status bclx_mlp_adam(bclx_handle* h, float* d_params, const float* d_grads, float* d_m, float* d_v,
                                   int t, float lr, float beta1, float beta2, float eps);

/* ---- SpMM: C[i*ldc + j] = beta * C + alpha * sum_k A[i,k] * B[k*ldb + j] for j < n -------------------------
 * A fp32 CSR (ptr_width 4 or 8), B row-major fp32 (b_dtype BCLX_F32) or bf16 (BCLX_BF16), C row-major fp32,
 * products accumulated in fp32. BCL's spmm_wrapper is alpha = beta = 1. */
BCLX_API bclx_status bclx_spmm(bclx_handle* h, const bclx_csr* A, const void* d_B, int32_t b_dtype, int64_t ldb,
                               float* d_C, int64_t ldc, int64_t n, float alpha, float beta);

/* The same for fp64: A fp64 CSR, B and C row-major fp64, accumulated in fp64 (the reference's spmm_wrapper on
 * double, spgemm.hpp:119-149). */
BCLX_API bclx_status bclx_spmm_f64(bclx_handle* h, const bclx_csr* A, const double* d_B, int64_t ldb, double* d_C,
                                   int64_t ldc, int64_t n, double alpha, double beta);

/* ---- dense GEMM: C = beta * C + alpha * op(A) * op(B), row-major (BCL calls it with alpha = beta = 1) --------
 * mode 0: fp32 inputs, FP32 MFMA (xf32-free, exact fp32 products); mode 1: inputs rounded to bf16, BF16 MFMA
 * with FP32 accumulation; mode 2: split-bf16 3-term (a_hi b_hi + a_hi b_lo + a_lo b_hi) on BF16 MFMA with FP32
 * accumulation. A, B, C fp32. trans_a / trans_b: 0 = no transpose, 1 = transpose. */
BCLX_API bclx_status bclx_gemm(bclx_handle* h, int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k,
                               float alpha, const float* d_A, int64_t lda, const float* d_B, int64_t ldb, float beta,
                               float* d_C, int64_t ldc, int32_t mode);
/* The same for fp64 operands (cblas_gemm_wrapper_'s double overload -> cblas_dgemm,
 * bcl/containers/algorithms/cblas_wrapper.hpp:29-40): C = beta * C + alpha * op(A) * op(B), row-major, fp64 MFMA. */
BCLX_API bclx_status bclx_gemm_f64(bclx_handle* h, int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k,
                                   double alpha, const double* d_A, int64_t lda, const double* d_B, int64_t ldb,
                                   double beta, double* d_C, int64_t ldc);
/* FP64 dense GEMM C = A * B + C (fp32 inputs widened exactly, fp64 C): the GPU error oracle for large sizes. */
BCLX_API bclx_status bclx_gemm_f64_ref(bclx_handle* h, int64_t m, int64_t n, int64_t k, const float* d_A,
                                       int64_t lda, const float* d_B, int64_t ldb, double* d_C, int64_t ldc);

/* ---- tile ingest and dense tile arithmetic ----------------------------------------------------------------
 * bclx_csr_split_grid: the tiles of A (device CSR, columns sorted within rows) on a grid of tile_m x tile_n tiles,
 * tile (ti, tj) = rows [ti tile_m, (ti+1) tile_m) and columns [tj tile_n, (tj+1) tile_n) with columns shifted to 0,
 * as SPMatrix's constructor cuts a matrix into tiles (SPMatrix::init_, bcl/containers/SPMatrix.hpp:212-266, through
 * CSRMatrix::get_slice_impl_, bcl/containers/sequential/CSRMatrix.hpp:238-276). tiles holds ntiles =
 * ceil(m / tile_m) * ceil(n / tile_n) descriptors, row-major over the grid. Two calls: with every tiles[t].row_ptr
 * NULL, each descriptor gets m, n, nnz and dtype; then, with tiles[t].row_ptr (m + 1 of tiles[t].ptr_width),
 * col_idx and vals (nnz each) allocated by the caller, the tiles are written. At most 256 tile columns. Columns that
 * are not sorted or not below A->n fail with BCLX_ERR_INVALID. */
BCLX_API bclx_status bclx_csr_split_grid(bclx_handle* h, const bclx_csr* A, int64_t tile_m, int64_t tile_n,
                                         int32_t ntiles, bclx_csr* tiles);
/* C = A op B elementwise on m x n row-major tiles (op 0 +, 1 -, 2 *, 3 /; dtype BCLX_F32 or BCLX_F64), each with its
 * own leading dimension; C may be A (in place, with ldc == lda). DMatrix::binary_op / binary_op_inplace and its
 * operators +, -, *, / and += (bcl/containers/DMatrix.hpp:319-386) on one tile. */
BCLX_API bclx_status bclx_dense_binary_op(bclx_handle* h, int32_t op, int32_t dtype, int64_t m, int64_t n,
                                          const void* d_A, int64_t lda, const void* d_B, int64_t ldb, void* d_C,
                                          int64_t ldc);

/* ---- MatrixMarket files (host memory, no GPU needed) -------------------------------------------------------
 * bclx_mm_info: m, n and the stored entry count from the size line
 *     (BCL::matrix_io::matrix_market_info, bcl/containers/sequential/matrix_io.hpp:68-98).
 * bclx_mm_read: a coordinate file (pattern / numeric field, general / symmetric) as a host CSR matrix sorted by
 *     (i, j), the first occurrence of a repeated (i, j) kept (BCL::matrix_io::mmread, matrix_io.hpp:221-357, and
 *     CSRMatrix::read_MatrixMarket, bcl/containers/sequential/CSRMatrix.hpp:369-411), in time linear in the file.
 *     out->row_ptr (m + 1 entries of ptr_width), out->col_idx and out->vals (dtype) are malloc'd host arrays owned by
 *     the caller (bclx_host_free). Errors (BCLX_ERR_INVALID) carry the reference's messages: not a Matrix Market
 *     file, unsupported matrix type, nonzero out of bounds, more nonzeros than reported.
 * bclx_mm_write: a host CSR matrix in the reference's output format (CSRMatrix::write_MatrixMarket,
 *     CSRMatrix.hpp:532-548): "integer general" banner, 1-based indices, values printed with "%f". */
BCLX_API bclx_status bclx_mm_info(const char* path, int32_t one_indexed, int64_t* m, int64_t* n, int64_t* nnz);
BCLX_API bclx_status bclx_mm_read(const char* path, int32_t one_indexed, int32_t dtype, int32_t ptr_width,
                                  bclx_csr* out);
BCLX_API bclx_status bclx_mm_write(const char* path, const bclx_csr* host_csr);
BCLX_API void bclx_host_free(void* p);

/* Library version string. */
BCLX_API const char* bclx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* BCLX_H */
