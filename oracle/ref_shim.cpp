// C entry points around the reference's own CPU tile multiply -- TEST INFRASTRUCTURE ONLY.
//
// This file holds no algorithm: it includes the reference headers from the reference source tree (the
// include path is set by oracle/Makefile) and calls
//   BCL::gemm(CSRMatrix, CSRMatrix, CSRMatrix)       bcl/containers/algorithms/spgemm.hpp:40-76
//   BCL::spmm_wrapper(...)                            bcl/containers/algorithms/spgemm.hpp:140-149
//   BCL::SparseHashAccumulator::get_matrix(m, n)     bcl/containers/sequential/SparseHashAccumulator.hpp:101-182
//   CSRMatrix::read_MatrixMarket / write_MatrixMarket, matrix_io::matrix_market_info,
//   CSRMatrix::get_slice_impl_ and CSRMatrixMemoryMapped::get_slice_impl_ (the tile ingest and file formats)
// so the C oracle can be checked bit for bit against the reference and the CPU baseline can time it.
// The headers pull in the MPI backend declarations (bcl/bcl.hpp:19-24); none of it runs on these paths, and
// the build drops the unused MPI-calling functions (-ffunction-sections, --gc-sections), so no MPI library
// is linked.
#include <bcl/bcl.hpp>
#include <bcl/containers/sequential/CSRMatrix.hpp>
#include <bcl/containers/algorithms/spgemm.hpp>
#include <bcl/containers/sequential/SparseHashAccumulator.hpp>
#include <bcl/containers/sequential/CSRMatrixMemoryMapped.hpp>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
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

thread_local std::string g_ref_error;

// CSRMatrix::read_MatrixMarket (CSRMatrix.hpp:369-411 over matrix_io::mmread, matrix_io.hpp:221-357). NULL on an
// exception, whose message ref_error() returns.
template <typename T>
void* mm_read(const char* path, int one_indexed) {
  try {
    BCL::CSRMatrix<T, int> a(0, 0);
    a.read_MatrixMarket(path, one_indexed != 0);
    return new RefResult<T>{(int64_t)a.m_, (int64_t)a.n_, (int64_t)a.nnz_, a.row_ptr_, a.col_ind_, a.vals_};
  } catch (const std::exception& e) {
    g_ref_error = e.what();
    return nullptr;
  }
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

REF_API void* ref_mm_read_f32(const char* path, int one_indexed) { return mm_read<float>(path, one_indexed); }
REF_API void* ref_mm_read_f64(const char* path, int one_indexed) { return mm_read<double>(path, one_indexed); }
REF_API const char* ref_error() { return g_ref_error.c_str(); }
REF_API int64_t ref_result_m(void* r, int is_f64) {
  return is_f64 ? static_cast<RefResult<double>*>(r)->m : static_cast<RefResult<float>*>(r)->m;
}
REF_API int64_t ref_result_n(void* r, int is_f64) {
  return is_f64 ? static_cast<RefResult<double>*>(r)->n : static_cast<RefResult<float>*>(r)->n;
}

// matrix_io::matrix_market_info (matrix_io.hpp:68-98): m, n, nnz of the size line.
REF_API void ref_mm_info(const char* path, int64_t* out) {
  auto s = BCL::matrix_io::matrix_market_info(path);
  out[0] = (int64_t)s.shape[0];
  out[1] = (int64_t)s.shape[1];
  out[2] = (int64_t)s.nnz;
}

// CSRMatrix::write_MatrixMarket (CSRMatrix.hpp:532-548) of a float/int matrix.
REF_API void ref_mm_write_f32(const char* path, int64_t m, int64_t n, int64_t nnz, const int* rp, const int* col,
                              const float* val) {
  auto a = make_csr<float>(m, n, nnz, rp, col, val);
  a.write_MatrixMarket(path);
}

// CSRMatrixMemoryMapped<float, int>(fname).get_slice_impl_(imin, imax, jmin, jmax)
// (bcl/containers/sequential/CSRMatrixMemoryMapped.hpp:55-145) of a binary CSR file.
REF_API void* ref_mmap_slice_f32(const char* path, int64_t imin, int64_t imax, int64_t jmin, int64_t jmax) {
  BCL::CSRMatrixMemoryMapped<float, int> f(path);
  auto c = f.get_slice_impl_(imin, imax, jmin, jmax);
  return new RefResult<float>{(int64_t)c.m_, (int64_t)c.n_, (int64_t)c.nnz_, c.row_ptr_, c.col_ind_, c.vals_};
}

// CSRMatrix::get_slice_impl_ (CSRMatrix.hpp:238-276) of an in-memory float/int matrix.
REF_API void* ref_slice_f32(int64_t m, int64_t n, int64_t nnz, const int* rp, const int* col, const float* val,
                            int64_t imin, int64_t imax, int64_t jmin, int64_t jmax) {
  auto a = make_csr<float>(m, n, nnz, rp, col, val);
  auto c = a.get_slice_impl_(imin, imax, jmin, jmax);
  return new RefResult<float>{(int64_t)c.m_, (int64_t)c.n_, (int64_t)c.nnz_, c.row_ptr_, c.col_ind_, c.vals_};
}

REF_API void ref_spmm_f32(int64_t m, int64_t n, int64_t k, int64_t nnz, float* c, float* b, float* a_val, int* a_rp,
                  int* a_col, int64_t ldb, int64_t ldc) {
  BCL::spmm_wrapper(m, n, k, nnz, c, b, a_val, a_rp, a_col, ldb, ldc);
}

}  // extern "C"
