/* CPU oracle for the local tile multiply under BCL::gemm -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a C restatement of the reference's CPU tile multiply, used by the tests, the smoke check and
 * the CPU-baseline leg of the benchmark as the checker. It is never the thing measured as the product
 * and never a fallback for the GPU path.
 *
 * Matrices are 0-based CSR with int64 row_ptr[m+1] and int32 col_idx; values fp32 or fp64. Functions
 * that produce a sparse result allocate it with malloc and hand ownership to the caller
 * (oracle_csr_free).
 *
 * Parity pins (checked in tests/test_oracle.py):
 *   oracle_spgemm_*        bit-exact (structure and fp32 value bits) against the reference's own
 *                          BCL::gemm(CSRMatrix, CSRMatrix, CSRMatrix) compiled from its headers
 *                          (oracle/_ref) and against the committed golden fixtures it produced;
 *                          structure also against scipy's boolean product (positive values), brute
 *                          force dense products on tiny inputs, A*I = A and permutations.
 *   oracle_spgemm_rowhash_*  same outputs as oracle_spgemm_* (structure bit-exact, values within
 *                          fp32 reordering), closed-form Laplacian^2.
 *   oracle_spmm_*          against the reference spmm_wrapper (oracle/_ref), B = ones closed form.
 *   oracle_gemm_f64        against numpy float64 matmul and closed-form fill_range inputs.
 *   oracle_accumulate_*    against the reference SparseHashAccumulator::get_matrix (oracle/_ref) and
 *                          the single-shot invariant sum_k A_ik B_kj = A_i: B_:j.
 */
#ifndef BCLX_ORACLE_H
#define BCLX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int64_t m, n, nnz;
  int64_t* row_ptr; /* m + 1 */
  int32_t* col_idx; /* nnz */
  void* vals;       /* nnz of float or double */
} oracle_csr;

void oracle_csr_free(oracle_csr* c);

/* BCL::gemm(CSR a, CSR b, CSR c): c += a * b, exactly as bcl/containers/algorithms/spgemm.hpp:40-76 followed by
 * CSRMatrix::assign (bcl/containers/sequential/CSRMatrix.hpp:106-165). c0 may be NULL (empty C). When
 * dead_row_alloc != 0 the per-row std::vector<bool>(ncols) of spgemm.hpp:59 is reproduced (timing fidelity). */
int oracle_spgemm_f32(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col, const float* a_val,
                      const int64_t* b_rp, const int32_t* b_col, const float* b_val, const oracle_csr* c0,
                      int dead_row_alloc, oracle_csr* out);
/* Same algorithm in float64 (the FP64 value oracle; fp32 inputs are widened exactly). */
int oracle_spgemm_f64(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col, const double* a_val,
                      const int64_t* b_rp, const int32_t* b_col, const double* b_val, const oracle_csr* c0,
                      int dead_row_alloc, oracle_csr* out);
int oracle_spgemm_f32_to_f64(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col,
                             const float* a_val, const int64_t* b_rp, const int32_t* b_col, const float* b_val,
                             oracle_csr* out);

/* Row-parallel hash SpGEMM (C = A * B): per row a SimpleHash table of pow2 >= max(16, products_i) slots probed
 * linearly from col * 107 (bcl/containers/sequential/SimpleHash.hpp:24-165, hash_functions.hpp:27-31), then a
 * per-row sort by column (SparseHashAccumulator.hpp:55-63). OpenMP over rows. acc64 != 0 accumulates fp32 in
 * double and rounds once. */
int oracle_spgemm_rowhash_f32(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col,
                              const float* a_val, const int64_t* b_rp, const int32_t* b_col, const float* b_val,
                              int nthreads, oracle_csr* out);
int oracle_spgemm_rowhash_f64(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col,
                              const double* a_val, const int64_t* b_rp, const int32_t* b_col, const double* b_val,
                              int nthreads, oracle_csr* out);

/* Symbolic only: nnz(C_i) per row (int64 out[m]) and products_i (int64 prod[m], may be NULL). */
int oracle_spgemm_row_nnz(int64_t m, int64_t k, int64_t n, const int64_t* a_rp, const int32_t* a_col,
                          const int64_t* b_rp, const int32_t* b_col, int nthreads, int64_t* out, int64_t* prod);

/* spmm_wrapper_generic (spgemm.hpp:78-96): C[i*ldc + j] += A[i,k] * B[k*ldb + j], alpha = beta = 1. */
void oracle_spmm_f32(int64_t m, int64_t n, int64_t k, const int64_t* a_rp, const int32_t* a_col, const float* a_val,
                     const float* b, int64_t ldb, float* c, int64_t ldc);
/* FP64 oracle of the same product: c64 = c0 + A * B with B given as fp32 (b) or bf16 bits (b_bf16); one of them. */
void oracle_spmm_f64(int64_t m, int64_t n, int64_t k, const int64_t* a_rp, const int32_t* a_col, const float* a_val,
                     const float* b, const uint16_t* b_bf16, int64_t ldb, const float* c0, int64_t ldc, double* c64,
                     int nthreads);

/* Dense C += A * B in float64, row-major, lda/ldb/ldc leading dimensions; A and B fp32 (widened exactly).
 * rows [r0, r1) of C only (row-sampled error checks). c is float64. */
void oracle_gemm_f64(int64_t m, int64_t n, int64_t k, const float* a, int64_t lda, const float* b, int64_t ldb,
                     double* c, int64_t ldc, int64_t r0, int64_t r1, int nthreads);

/* SparseHashAccumulator::get_matrix (bcl/containers/sequential/SparseHashAccumulator.hpp:101-182): sum of ntiles
 * partial C tiles (all m x n, offsets 0), summed in tile order per entry, rows sorted by column. */
int oracle_accumulate_f32(int64_t m, int64_t n, int ntiles, const int64_t* const* rps, const int32_t* const* cols,
                          const float* const* vals, int nthreads, oracle_csr* out);

/* MLP 16 -> 64 -> 64 -> 1 (ReLU, ReLU, linear) in float64: forward for nrows feature rows and, when grads != NULL,
 * the gradient of the mean squared error 0.5/N * sum (y - t)^2 over all parameters.
 * Parameter layout: W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3[1] (5313 values). */
void oracle_mlp_forward_backward(int64_t nrows, const float* x, const float* target, const float* params, double* y,
                                 double* grads, double* loss);

#ifdef __cplusplus
}
#endif
#endif
