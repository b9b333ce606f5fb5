// C entry points around the reference's own CPU tile multiply -- TEST INFRASTRUCTURE ONLY.
//
// This file holds no algorithm: it includes the reference headers from the reference source tree (the
// include path is set by oracle/Makefile) and calls
//   BCL::gemm(CSRMatrix, CSRMatrix, CSRMatrix)       bcl/containers/algorithms/spgemm.hpp:40-76
//   BCL::spmm_wrapper(...)                            bcl/containers/algorithms/spgemm.hpp:140-149
//   BCL::SparseHashAccumulator::get_matrix(m, n)     bcl/containers/sequential/SparseHashAccumulator.hpp:101-182
// so the C oracle can be checked bit for bit against the reference and the CPU baseline can time it.
// The headers pull in the MPI backend declarations (bcl/bcl.hpp:19-24); none of it runs on these paths, and
// the build drops the unused MPI-calling functions (-ffunction-sections, --gc-sections), so no MPI library
// is linked.
#include <bcl/bcl.hpp>
#include <bcl/containers/sequential/CSRMatrix.hpp>
#include <bcl/containers/algorithms/spgemm.hpp>
#include <bcl/containers/sequential/SparseHashAccumulator.hpp>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

template <typename T>
struct RefResult {
  int64_t m, n, nnz;
  std::vector<int> row_ptr, col_ind;
  std::vector<T> vals;
};

template <typename T>
BCL::CSRMatrix<T, int> make_csr(int64_t m, int64_t n, int64_t nnz, const int* rp, const int* col, const T* val) {
  std::vector<T> v(val, val + nnz);
  std::vector<int> r(rp, rp + m + 1), c(col, col + nnz);
  return BCL::CSRMatrix<T, int>(m, n, nnz, std::move(v), std::move(r), std::move(c));
}

template <typename T>
void* spgemm(int64_t m, int64_t k, int64_t n, int64_t a_nnz, const int* a_rp, const int* a_col, const T* a_val,
             int64_t b_nnz, const int* b_rp, const int* b_col, const T* b_val, int64_t c_nnz, const int* c_rp,
             const int* c_col, const T* c_val) {
  auto a = make_csr<T>(m, k, a_nnz, a_rp, a_col, a_val);
  auto b = make_csr<T>(k, n, b_nnz, b_rp, b_col, b_val);
  BCL::CSRMatrix<T, int> c(m, n);
  if (c_rp) c = make_csr<T>(m, n, c_nnz, c_rp, c_col, c_val);
  BCL::gemm(a, b, c);  // c += a * b (spgemm.hpp:40-76)
  auto* r = new RefResult<T>{(int64_t)c.m_, (int64_t)c.n_, (int64_t)c.nnz_, c.row_ptr_, c.col_ind_, c.vals_};
  return r;
}

template <typename T>
void* accumulate(int64_t m, int64_t n, int ntiles, const int64_t* nnzs, const int* const* rps,
                 const int* const* cols, const T* const* vals) {
  BCL::SparseHashAccumulator<T, int> acc;
  for (int t = 0; t < ntiles; t++) acc.accumulate(make_csr<T>(m, n, nnzs[t], rps[t], cols[t], vals[t]));
  auto c = acc.get_matrix(m, n);
  auto* r = new RefResult<T>{(int64_t)c.m_, (int64_t)c.n_, (int64_t)c.nnz_, c.row_ptr_, c.col_ind_, c.vals_};
  return r;
}

}  // namespace

#define REF_API __attribute__((visibility("default")))

extern "C" {

REF_API void* ref_spgemm_f32(int64_t m, int64_t k, int64_t n, int64_t a_nnz, const int* a_rp, const int* a_col,
                     const float* a_val, int64_t b_nnz, const int* b_rp, const int* b_col, const float* b_val,
                     int64_t c_nnz, const int* c_rp, const int* c_col, const float* c_val) {
  return spgemm<float>(m, k, n, a_nnz, a_rp, a_col, a_val, b_nnz, b_rp, b_col, b_val, c_nnz, c_rp, c_col, c_val);
}

REF_API void* ref_spgemm_f64(int64_t m, int64_t k, int64_t n, int64_t a_nnz, const int* a_rp, const int* a_col,
                     const double* a_val, int64_t b_nnz, const int* b_rp, const int* b_col, const double* b_val,
                     int64_t c_nnz, const int* c_rp, const int* c_col, const double* c_val) {
  return spgemm<double>(m, k, n, a_nnz, a_rp, a_col, a_val, b_nnz, b_rp, b_col, b_val, c_nnz, c_rp, c_col, c_val);
}

REF_API void* ref_accumulate_f32(int64_t m, int64_t n, int ntiles, const int64_t* nnzs, const int* const* rps,
                         const int* const* cols, const float* const* vals) {
  return accumulate<float>(m, n, ntiles, nnzs, rps, cols, vals);
}

// Result accessors: sizes, then copy out and free.
REF_API int64_t ref_result_nnz(void* r, int is_f64) {
  return is_f64 ? static_cast<RefResult<double>*>(r)->nnz : static_cast<RefResult<float>*>(r)->nnz;
}

REF_API void ref_result_copy_free(void* r, int is_f64, int* row_ptr, int* col_ind, void* vals) {
  if (is_f64) {
    auto* x = static_cast<RefResult<double>*>(r);
    std::memcpy(row_ptr, x->row_ptr.data(), sizeof(int) * x->row_ptr.size());
    std::memcpy(col_ind, x->col_ind.data(), sizeof(int) * x->nnz);
    std::memcpy(vals, x->vals.data(), sizeof(double) * x->nnz);
    delete x;
  } else {
    auto* x = static_cast<RefResult<float>*>(r);
    std::memcpy(row_ptr, x->row_ptr.data(), sizeof(int) * x->row_ptr.size());
    std::memcpy(col_ind, x->col_ind.data(), sizeof(int) * x->nnz);
    std::memcpy(vals, x->vals.data(), sizeof(float) * x->nnz);
    delete x;
  }
}

// CSRMatrix::read_Binary then write_Binary (bcl/containers/sequential/CSRMatrix.hpp:488-530) on a float/int file.
REF_API int64_t ref_binary_roundtrip(const char* in, const char* out) {
  BCL::CSRMatrix<float, int> a(0, 0);
  a.read_Binary(in);
  a.write_Binary(out);
  return (int64_t)a.nnz_;
}

REF_API void ref_spmm_f32(int64_t m, int64_t n, int64_t k, int64_t nnz, float* c, float* b, float* a_val, int* a_rp,
                  int* a_col, int64_t ldb, int64_t ldc) {
  BCL::spmm_wrapper(m, n, k, nnz, c, b, a_val, a_rp, a_col, ldb, ldc);
}

}  // extern "C"
