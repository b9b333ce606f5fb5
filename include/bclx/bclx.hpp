// C++ host layer over the bclx C ABI: the reference's local tile multiply interfaces with the same names, argument
// meaning and error behaviour (std::runtime_error), running on the GPU.
//
//   bclx::CSRMatrix<T, I>        BCL::CSRMatrix            bcl/containers/sequential/CSRMatrix.hpp:32-366
//                                (same public members m_, n_, nnz_, vals_, row_ptr_, col_ind_; binary file format
//                                of read_Binary / write_Binary, CSRMatrix.hpp:488-530; get_slice_impl_ :238-276)
//   bclx::gemm(a, b, c)          BCL::gemm(CSRMatrix, CSRMatrix, CSRMatrix)  bcl/containers/algorithms/spgemm.hpp:40-76
//                                (c += a * b; exact-zero sums are kept)
//   a.dot(b)                     CSRMatrix::dot            CSRMatrix.hpp:229-235
//   bclx::spmm_wrapper(...)      BCL::spmm_wrapper         spgemm.hpp:140-149 (c += a * b, row-major, ldb / ldc)
//   bclx::cblas_gemm_wrapper_()  BCL::experimental::cblas_gemm_wrapper_  cblas_wrapper.hpp:16-40
//   bclx::SparseHashAccumulator  BCL::SparseHashAccumulator  SparseHashAccumulator.hpp:86-183
//
// Host containers hold std::vector data like the reference; each call copies operands to HBM, runs the HIP kernels
// through the C ABI on a process-wide handle (device 0 unless bclx::set_device is called first, default stream) and
// copies results back. Values of sparse products are accumulated in float64 and rounded once.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../bclx.h"

namespace bclx {

// CBLAS enumerator values (CBLAS_ORDER / CBLAS_TRANSPOSE), so CBLAS constants cast to int can be passed through.
enum Layout { RowMajor = 101, ColMajor = 102 };
enum Transpose { NoTrans = 111, Trans = 112, ConjTrans = 113 };
enum GemmMode { GemmFP32 = 0, GemmBF16 = 1, GemmSplitBF16 = 2 };

namespace detail {

inline void check(bclx_status s) {
  if (s != BCLX_OK) throw std::runtime_error(std::string("bclx: ") + bclx_last_error());
}
inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("bclx: ") + what + ": " + hipGetErrorString(e));
}

struct Context {
  int device = 0;
  bclx_handle* h = nullptr;
  GemmMode gemm_mode = GemmFP32;
  ~Context() {
    if (h) bclx_destroy(h);
  }
  bclx_handle* handle() {
    if (!h) check(bclx_create(device, nullptr, &h));
    return h;
  }
};
inline Context& ctx() {
  static Context c;
  return c;
}

// Device copy of a host array (RAII).
template <typename T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  explicit Dev(size_t count) : n(count) {
    hip_check(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * std::max<size_t>(count, 1)), "hipMalloc");
  }
  Dev(const T* host, size_t count) : Dev(count) {
    if (count) hip_check(hipMemcpy(p, host, sizeof(T) * count, hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  void to_host(T* host, size_t count) const {
    if (count) hip_check(hipMemcpy(host, p, sizeof(T) * count, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  }
};

template <typename T>
constexpr int32_t dtype_of() {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "values must be float or double");
  return std::is_same<T, float>::value ? BCLX_F32 : BCLX_F64;
}
template <typename I>
constexpr int32_t width_of() {
  static_assert(std::is_integral<I>::value && (sizeof(I) == 4 || sizeof(I) == 8), "index type must be 32 or 64 bit");
  return (int32_t)sizeof(I);
}

}  // namespace detail

inline void set_device(int device) {
  auto& c = detail::ctx();
  if (c.h) {
    detail::check(bclx_destroy(c.h));
    c.h = nullptr;
  }
  c.device = device;
}
inline void set_gemm_mode(GemmMode m) { detail::ctx().gemm_mode = m; }

template <typename T, typename I = int>
struct CSRMatrix {
  using size_type = size_t;
  using value_type = T;
  using index_type = I;

  size_type m_ = 0, n_ = 0, nnz_ = 0;
  std::vector<T> vals_;
  std::vector<I> row_ptr_;
  std::vector<I> col_ind_;

  CSRMatrix() = default;
  CSRMatrix(size_type m, size_type n) : m_(m), n_(n), nnz_(0), row_ptr_(m + 1, 0) {}
  CSRMatrix(size_type m, size_type n, size_type nnz) : m_(m), n_(n), nnz_(nnz), vals_(nnz), row_ptr_(m + 1), col_ind_(nnz) {}
  CSRMatrix(size_type m, size_type n, size_type nnz, std::vector<T>&& vals, std::vector<I>&& row_ptr,
            std::vector<I>&& col_ind)
      : m_(m), n_(n), nnz_(nnz), vals_(std::move(vals)), row_ptr_(std::move(row_ptr)), col_ind_(std::move(col_ind)) {}
  explicit CSRMatrix(const std::string& fname) { read_Binary(fname); }

  size_type m() const { return m_; }
  size_type n() const { return n_; }
  size_type nnz() const { return nnz_; }
  std::array<size_type, 2> shape() const noexcept { return {m_, n_}; }
  T* values_data() { return vals_.data(); }
  I* rowptr_data() { return row_ptr_.data(); }
  I* colind_data() { return col_ind_.data(); }

  bool operator==(const CSRMatrix& o) const {
    return m_ == o.m_ && n_ == o.n_ && nnz_ == o.nnz_ && vals_ == o.vals_ && row_ptr_ == o.row_ptr_ &&
           col_ind_ == o.col_ind_;
  }

  // Build from ((i, j), v) tuples in any order (CSRMatrix::assign): sorted by (i, j); duplicates are not merged.
  template <typename InputIt>
  void assign(InputIt first, InputIt last) {
    std::vector<std::pair<std::pair<I, I>, T>> t(first, last);
    std::sort(t.begin(), t.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    nnz_ = t.size();
    vals_.resize(nnz_);
    col_ind_.resize(nnz_);
    row_ptr_.assign(m_ + 1, 0);
    for (size_type c = 0; c < nnz_; c++) {
      vals_[c] = t[c].second;
      col_ind_[c] = t[c].first.second;
      row_ptr_[t[c].first.first + 1]++;
    }
    for (size_type r = 0; r < m_; r++) row_ptr_[r + 1] += row_ptr_[r];
  }

  // this * other on the GPU (a fresh result, as CSRMatrix::dot)
  CSRMatrix dot(const CSRMatrix& other) const;

  CSRMatrix get_slice_impl_(size_type imin, size_type imax, size_type jmin, size_type jmax) const {
    imax = std::min(imax, m_);
    jmax = std::min(jmax, n_);
    CSRMatrix r(imax - imin, jmax - jmin);
    for (size_type i = imin; i < imax; i++) {
      for (I p = row_ptr_[i]; p < row_ptr_[i + 1]; p++) {
        const size_type j = (size_type)col_ind_[p];
        if (j >= jmin && j < jmax) {
          r.vals_.push_back(vals_[p]);
          r.col_ind_.push_back((I)(j - jmin));
        }
      }
      r.row_ptr_[i - imin + 1] = (I)r.vals_.size();
    }
    r.nnz_ = r.vals_.size();
    return r;
  }

  // Binary CSR file: m, n, nnz as size_t, then vals (T), col_ind (I), row_ptr (I).
  void read_Binary(const std::string& fname) {
    FILE* f = fopen(fname.c_str(), "rb");
    if (!f) throw std::runtime_error("bclx: cannot open " + fname);
    bool ok = fread(&m_, sizeof(size_t), 1, f) == 1 && fread(&n_, sizeof(size_t), 1, f) == 1 &&
              fread(&nnz_, sizeof(size_t), 1, f) == 1;
    vals_.resize(nnz_);
    col_ind_.resize(nnz_);
    row_ptr_.resize(m_ + 1);
    ok = ok && fread(vals_.data(), sizeof(T), nnz_, f) == nnz_ && fread(col_ind_.data(), sizeof(I), nnz_, f) == nnz_ &&
         fread(row_ptr_.data(), sizeof(I), m_ + 1, f) == m_ + 1;
    fclose(f);
    if (!ok) throw std::runtime_error("bclx: short read in " + fname);
  }
  void write_Binary(const std::string& fname) const {
    FILE* f = fopen(fname.c_str(), "wb");
    if (!f) throw std::runtime_error("bclx: cannot open " + fname);
    fwrite(&m_, sizeof(size_t), 1, f);
    fwrite(&n_, sizeof(size_t), 1, f);
    fwrite(&nnz_, sizeof(size_t), 1, f);
    fwrite(vals_.data(), sizeof(T), nnz_, f);
    fwrite(col_ind_.data(), sizeof(I), nnz_, f);
    fwrite(row_ptr_.data(), sizeof(I), m_ + 1, f);
    fclose(f);
  }
};

namespace detail {

// Device-resident copy of a host CSRMatrix with its C-ABI descriptor.
template <typename T, typename I>
struct DevCSR {
  Dev<I> rp;
  Dev<int32_t> col;
  Dev<T> val;
  bclx_csr d{};
  explicit DevCSR(const CSRMatrix<T, I>& a)
      : rp(a.row_ptr_.data(), a.m_ + 1), col(a.nnz_), val(a.vals_.data(), a.nnz_) {
    std::vector<int32_t> c(a.col_ind_.begin(), a.col_ind_.end());
    if (a.nnz_) hip_check(hipMemcpy(col.p, c.data(), 4 * a.nnz_, hipMemcpyHostToDevice), "hipMemcpy H2D");
    d = bclx_csr{(int64_t)a.m_, (int64_t)a.n_, (int64_t)a.nnz_, width_of<I>(), dtype_of<T>(), rp.p, col.p, val.p};
  }
};

template <typename T, typename I>
CSRMatrix<T, I> to_host(const Dev<I>& rp, const Dev<int32_t>& col, const Dev<T>& val, size_t m, size_t n,
                        size_t nnz) {
  CSRMatrix<T, I> c(m, n, nnz);
  rp.to_host(c.row_ptr_.data(), m + 1);
  std::vector<int32_t> cc(nnz);
  col.to_host(cc.data(), nnz);
  std::copy(cc.begin(), cc.end(), c.col_ind_.begin());
  val.to_host(c.vals_.data(), nnz);
  return c;
}

template <typename T, typename I>
CSRMatrix<T, I> multiply(const CSRMatrix<T, I>& a, const CSRMatrix<T, I>& b) {
  if (a.n_ != b.m_) throw std::runtime_error("bclx: inner dimensions differ");
  bclx_handle* h = ctx().handle();
  DevCSR<T, I> A(a), B(b);
  bclx_spgemm_plan* plan = nullptr;
  int64_t nnz = 0;
  check(bclx_spgemm_symbolic(h, &A.d, &B.d, &plan, &nnz));
  std::unique_ptr<bclx_spgemm_plan, bclx_status (*)(bclx_spgemm_plan*)> guard(plan, bclx_spgemm_plan_destroy);
  Dev<I> rp(a.m_ + 1);
  Dev<int32_t> col((size_t)nnz);
  Dev<T> val((size_t)nnz);
  bclx_csr C{(int64_t)a.m_, (int64_t)b.n_, nnz, width_of<I>(), dtype_of<T>(), rp.p, col.p, val.p};
  check(bclx_spgemm_numeric(plan, &C));
  check(bclx_synchronize(h));
  return to_host<T, I>(rp, col, val, a.m_, b.n_, (size_t)nnz);
}

template <typename T, typename I>
CSRMatrix<T, I> sum(const std::vector<const CSRMatrix<T, I>*>& mats, size_t m, size_t n) {
  bclx_handle* h = ctx().handle();
  std::vector<std::unique_ptr<DevCSR<T, I>>> dev;
  std::vector<bclx_csr> descs;
  for (auto* x : mats) {
    if (x->m_ != m || x->n_ != n) throw std::runtime_error("bclx: accumulated tiles must be m x n");
    dev.push_back(std::make_unique<DevCSR<T, I>>(*x));
    descs.push_back(dev.back()->d);
  }
  Dev<I> rp(m + 1);
  bclx_csr C{(int64_t)m, (int64_t)n, 0, width_of<I>(), dtype_of<T>(), rp.p, nullptr, nullptr};
  check(bclx_accumulate_tiles(h, (int)descs.size(), descs.data(), &C));
  const int64_t nnz = C.nnz;
  Dev<int32_t> col((size_t)nnz);
  Dev<T> val((size_t)nnz);
  if (nnz > 0) {
    C.col_idx = col.p;
    C.vals = val.p;
    check(bclx_accumulate_tiles(h, (int)descs.size(), descs.data(), &C));
  }
  check(bclx_synchronize(h));
  return to_host<T, I>(rp, col, val, m, n, (size_t)nnz);
}

}  // namespace detail

template <typename T, typename I>
CSRMatrix<T, I> CSRMatrix<T, I>::dot(const CSRMatrix& other) const {
  return detail::multiply(*this, other);
}

// c += a * b (BCL::gemm(CSRMatrix, CSRMatrix, CSRMatrix)): the existing entries of c are kept and summed with the
// product; new entries are created for every product position, including exact-zero sums.
template <typename T, typename I>
void gemm(const CSRMatrix<T, I>& a, const CSRMatrix<T, I>& b, CSRMatrix<T, I>& c) {
  if (c.m_ != a.m_ || c.n_ != b.n_) throw std::runtime_error("bclx: c has the wrong shape");
  CSRMatrix<T, I> p = detail::multiply(a, b);
  if (c.nnz_ == 0) {
    c = std::move(p);
    return;
  }
  c = detail::sum<T, I>({&c, &p}, c.m_, c.n_);
}

// BCL::SparseHashAccumulator: collect partial C tiles, then get_matrix(m, n) sums them (on the GPU).
template <typename T, typename I = int>
struct SparseHashAccumulator {
  std::vector<CSRMatrix<T, I>> mats_;
  void accumulate(CSRMatrix<T, I>&& mat) { mats_.push_back(std::move(mat)); }
  CSRMatrix<T, I> get_matrix(size_t m, size_t n) const {
    if (mats_.empty()) return CSRMatrix<T, I>(m, n);
    std::vector<const CSRMatrix<T, I>*> p;
    for (auto& x : mats_) p.push_back(&x);
    return detail::sum<T, I>(p, m, n);
  }
};

// BCL::spmm_wrapper: c[i*ldc + j] += sum_k a[i,k] * b[k*ldb + j], j < n; host arrays, fp32 values.
template <typename I>
void spmm_wrapper(size_t m, size_t n, size_t k, size_t nnz, float* c, float* b, float* a_values, I* a_rowptr,
                  I* a_colind, size_t ldb, size_t ldc) {
  bclx_handle* h = detail::ctx().handle();
  detail::Dev<I> rp(a_rowptr, m + 1);
  std::vector<int32_t> ci(a_colind, a_colind + nnz);
  detail::Dev<int32_t> col(ci.data(), nnz);
  detail::Dev<float> val(a_values, nnz);
  detail::Dev<float> B(b, k * ldb), C(c, m * ldc);
  bclx_csr A{(int64_t)m, (int64_t)k, (int64_t)nnz, detail::width_of<I>(), BCLX_F32, rp.p, col.p, val.p};
  detail::check(bclx_spmm(h, &A, B.p, BCLX_F32, (int64_t)ldb, C.p, (int64_t)ldc, (int64_t)n, 1.f, 1.f));
  detail::check(bclx_synchronize(h));
  C.to_host(c, m * ldc);
}

// BCL::experimental::cblas_gemm_wrapper_ (float): c = beta * c + alpha * op(a) * op(b) on the matrix cores,
// arithmetic mode from set_gemm_mode (FP32 MFMA by default).
inline void cblas_gemm_wrapper_(int layout, int transa, int transb, int m, int n, int k, float alpha, const float* a,
                                int lda, const float* b, int ldb, float beta, float* c, int ldc) {
  bclx_handle* h = detail::ctx().handle();
  const bool ta = transa != NoTrans, tb = transb != NoTrans;
  if (layout == ColMajor) {
    // column-major C = op(A) op(B) is row-major C^T = op(B)^T op(A)^T: swap the operands and the sizes
    std::swap(m, n);
    std::swap(a, b);
    std::swap(lda, ldb);
    const bool t = ta;
    return cblas_gemm_wrapper_(RowMajor, tb ? Trans : NoTrans, t ? Trans : NoTrans, m, n, k, alpha, a, lda, b, ldb,
                               beta, c, ldc);
  }
  const size_t a_rows = ta ? k : m, b_rows = tb ? n : k;
  detail::Dev<float> A(a, a_rows * (size_t)lda), B(b, b_rows * (size_t)ldb), C(c, (size_t)m * ldc);
  detail::check(bclx_gemm(h, ta, tb, m, n, k, alpha, A.p, lda, B.p, ldb, beta, C.p, ldc, detail::ctx().gemm_mode));
  detail::check(bclx_synchronize(h));
  C.to_host(c, (size_t)m * ldc);
}

}  // namespace bclx
