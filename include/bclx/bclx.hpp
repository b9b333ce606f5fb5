// C++ host layer over the bclx C ABI: the reference's local tile multiply interfaces with the same names, argument
// meaning and error behaviour (std::runtime_error), running on the GPU.
//
//   bclx::CSRMatrix<T, I, A>     BCL::CSRMatrix            bcl/containers/sequential/CSRMatrix.hpp:32-366
//                                (same public members m_, n_, nnz_, vals_, row_ptr_, col_ind_; binary file format
//                                of read_Binary / write_Binary, CSRMatrix.hpp:488-530; get_slice_impl_ :238-276)
//   bclx::gemm(a, b, c)          BCL::gemm(CSRMatrix, CSRMatrix, CSRMatrix)  bcl/containers/algorithms/spgemm.hpp:40-76
//                                (c += a * b; exact-zero sums are kept)
//   a.dot(b), bclx::dot(a, b)    CSRMatrix::dot            CSRMatrix.hpp:229-235
//   (bclx::dot, bclx::gemm, bclx::split_grid and bclx::SparseHashAccumulator also take BCL::CSRMatrix<T, I, A>
//   itself, or any type with its public members and constructor, and return that type)
//   bclx::spmm_wrapper(...)      BCL::spmm_wrapper         spgemm.hpp:140-149 (c += a * b, row-major, ldb / ldc;
//                                float or double)
//   bclx::cblas_gemm_wrapper_()  BCL::experimental::cblas_gemm_wrapper_  cblas_wrapper.hpp:16-40 (float and double)
//   bclx::SparseHashAccumulator  BCL::SparseHashAccumulator  SparseHashAccumulator.hpp:86-183
//   CSRMatrix::read_MatrixMarket / write_MatrixMarket, CSRMatrix(fname, format), matrix_io::matrix_info
//                                CSRMatrix.hpp:191-224, :369-411, :532-548; matrix_io.hpp:10-98, :221-357
//                                (read in linear time by libbclx; the reference's reader is quadratic)
//   bclx::CSRMatrixMemoryMapped  BCL::CSRMatrixMemoryMapped  CSRMatrixMemoryMapped.hpp:55-145
//   bclx::split_grid(a, tm, tn)  the tile cut of SPMatrix's constructor (SPMatrix.hpp:212-266), on the GPU
//   bclx::tile_binary_op(...)    DMatrix::binary_op / binary_op_inplace on one tile (DMatrix.hpp:319-386), on the GPU
//
// Host containers hold std::vector data like the reference; each call copies operands to HBM, runs the HIP kernels
// through the C ABI on a process-wide handle (device 0 unless bclx::set_device is called first, default stream) and
// copies results back. Values of sparse products are accumulated in float64 and rounded once.
#pragma once

#include <fcntl.h>
#include <hip/hip_runtime_api.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

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

// File formats of CSRMatrix(fname, format) (BCL::FileFormat, bcl/containers/sequential/matrix_io.hpp:10-15).
enum FileFormat { MatrixMarket, MatrixMarketZeroIndexed, Binary, Unknown };

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

// BCL::CSRMatrix<T, I, Allocator> (CSRMatrix.hpp:32-53): the same template parameters, member names and layout; the
// index arrays use the allocator rebound to I, as the reference's IAllocator.
template <typename T, typename I = int, typename Allocator = std::allocator<T>>
struct CSRMatrix {
  using size_type = size_t;
  using value_type = T;
  using index_t = I;
  using index_type = I;
  using allocator_traits = std::allocator_traits<Allocator>;
  using IAllocator = typename allocator_traits::template rebind_alloc<I>;

  size_type m_ = 0, n_ = 0, nnz_ = 0;
  std::vector<T, Allocator> vals_;
  std::vector<I, IAllocator> row_ptr_;
  std::vector<I, IAllocator> col_ind_;

  CSRMatrix() = default;
  CSRMatrix(size_type m, size_type n) : m_(m), n_(n), nnz_(0), row_ptr_(m + 1, 0) {}
  CSRMatrix(size_type m, size_type n, size_type nnz) : m_(m), n_(n), nnz_(nnz), vals_(nnz), row_ptr_(m + 1), col_ind_(nnz) {}
  CSRMatrix(size_type m, size_type n, size_type nnz, std::vector<T, Allocator>&& vals,
            std::vector<I, IAllocator>&& row_ptr, std::vector<I, IAllocator>&& col_ind)
      : m_(m), n_(n), nnz_(nnz), vals_(std::move(vals)), row_ptr_(std::move(row_ptr)), col_ind_(std::move(col_ind)) {}
  // From a file: ".mtx" is MatrixMarket, ".binary" BCL's binary format, unless the format is given
  // (CSRMatrix.hpp:212-224, matrix_io::detect_file_type).
  explicit CSRMatrix(const std::string& fname, FileFormat format = Unknown);

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

  // MatrixMarket coordinate file (1-based unless one_indexed is false), through libbclx's linear-time reader.
  void read_MatrixMarket(const std::string& fname, bool one_indexed = true);
  // The reference's output format: "integer general" banner, 1-based indices, values printed with "%f".
  void write_MatrixMarket(const std::string& fname) const;

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
  template <class M>  // any container with BCL::CSRMatrix's public members
  explicit DevCSR(const M& a)
      : rp(a.row_ptr_.data(), a.m_ + 1), col(a.nnz_), val(a.vals_.data(), a.nnz_) {
    std::vector<int32_t> c(a.col_ind_.begin(), a.col_ind_.end());
    if (a.nnz_) hip_check(hipMemcpy(col.p, c.data(), 4 * a.nnz_, hipMemcpyHostToDevice), "hipMemcpy H2D");
    d = bclx_csr{(int64_t)a.m_, (int64_t)a.n_, (int64_t)a.nnz_, width_of<I>(), dtype_of<T>(), rp.p, col.p, val.p};
  }
};

// A host matrix of type M (BCL::CSRMatrix's public members and its (m, n, nnz, vals, row_ptr, col_ind) constructor)
// from device arrays.
template <class M, typename T, typename I>
M to_host(const Dev<I>& rp, const Dev<int32_t>& col, const Dev<T>& val, size_t m, size_t n, size_t nnz) {
  decltype(M::vals_) v(nnz);
  decltype(M::row_ptr_) r(m + 1);
  decltype(M::col_ind_) c(nnz);
  rp.to_host(r.data(), m + 1);
  std::vector<int32_t> cc(nnz);
  col.to_host(cc.data(), nnz);
  std::copy(cc.begin(), cc.end(), c.begin());
  val.to_host(v.data(), nnz);
  return M(m, n, nnz, std::move(v), std::move(r), std::move(c));
}

template <class M>
M multiply(const M& a, const M& b) {
  using T = typename M::value_type;
  using I = typename M::index_type;
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
  return to_host<M>(rp, col, val, a.m_, b.n_, (size_t)nnz);
}

template <class M>
M sum(const std::vector<const M*>& mats, size_t m, size_t n) {
  using T = typename M::value_type;
  using I = typename M::index_type;
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
  return to_host<M>(rp, col, val, m, n, (size_t)nnz);
}

// Containers shaped like BCL::CSRMatrix (public vals_ / row_ptr_ / col_ind_ and value_type / index_type).
template <class M, class = void>
struct is_csr : std::false_type {};
template <class M>
struct is_csr<M, std::void_t<typename M::value_type, typename M::index_type, decltype(std::declval<M&>().vals_),
                             decltype(std::declval<M&>().row_ptr_), decltype(std::declval<M&>().col_ind_)>>
    : std::true_type {};

}  // namespace detail

template <typename T, typename I, typename A>
CSRMatrix<T, I, A> CSRMatrix<T, I, A>::dot(const CSRMatrix& other) const {
  return detail::multiply(*this, other);
}

// a * b on the GPU for any matrix type shaped like BCL::CSRMatrix<T, I, Allocator>, returned as that same type: the
// call `my_a.dot(my_b)` of BCL::gemm (spgemm.hpp:221) becomes `bclx::dot(my_a, my_b)` on BCL's own tiles, with no
// conversion of the operands or the result.
template <class M, typename = std::enable_if_t<detail::is_csr<M>::value>>
M dot(const M& a, const M& b) {
  return detail::multiply(a, b);
}

namespace matrix_io {

inline FileFormat detect_file_type(const std::string& fname) {
  const size_t dot = fname.rfind('.');
  const std::string suffix = dot == std::string::npos ? std::string() : fname.substr(dot);
  if (suffix == ".mtx") return MatrixMarket;
  if (suffix == ".binary") return Binary;
  return Unknown;
}

struct matrix_shape {
  std::array<size_t, 2> shape;
  size_t nnz;
};

// m, n and the stored entry count of a MatrixMarket or binary file (matrix_io::matrix_info, matrix_io.hpp:50-113)
inline matrix_shape matrix_info(const std::string& fname, FileFormat format = Unknown) {
  if (format == Unknown) format = detect_file_type(fname);
  if (format == MatrixMarket || format == MatrixMarketZeroIndexed) {
    int64_t m = 0, n = 0, nnz = 0;
    detail::check(bclx_mm_info(fname.c_str(), format == MatrixMarket, &m, &n, &nnz));
    return {{(size_t)m, (size_t)n}, (size_t)nnz};
  }
  if (format == Binary) {
    FILE* f = fopen(fname.c_str(), "rb");
    if (!f) throw std::runtime_error("bclx: cannot open " + fname);
    size_t h[3] = {0, 0, 0};
    const bool ok = fread(h, sizeof(size_t), 3, f) == 3;
    fclose(f);
    if (!ok) throw std::runtime_error("bclx: short read in " + fname);
    return {{h[0], h[1]}, h[2]};
  }
  throw std::runtime_error("matrix_info: Could not detect file format for \"" + fname + "\"");
}

}  // namespace matrix_io

template <typename T, typename I, typename A>
CSRMatrix<T, I, A>::CSRMatrix(const std::string& fname, FileFormat format) {
  if (format == Unknown) format = matrix_io::detect_file_type(fname);
  if (format == MatrixMarket) read_MatrixMarket(fname);
  else if (format == MatrixMarketZeroIndexed) read_MatrixMarket(fname, false);
  else if (format == Binary) read_Binary(fname);
  else throw std::runtime_error("CSRMatrix: Could not detect file format for \"" + fname + "\"");
}

template <typename T, typename I, typename A>
void CSRMatrix<T, I, A>::read_MatrixMarket(const std::string& fname, bool one_indexed) {
  bclx_csr c{};
  detail::check(bclx_mm_read(fname.c_str(), one_indexed, detail::dtype_of<T>(), 8, &c));
  std::unique_ptr<void, void (*)(void*)> rp(c.row_ptr, bclx_host_free), col(c.col_idx, bclx_host_free),
      val(c.vals, bclx_host_free);
  m_ = (size_type)c.m;
  n_ = (size_type)c.n;
  nnz_ = (size_type)c.nnz;
  const int64_t* r = static_cast<const int64_t*>(c.row_ptr);
  row_ptr_.assign(r, r + m_ + 1);
  col_ind_.assign(c.col_idx, c.col_idx + nnz_);
  const T* v = static_cast<const T*>(c.vals);
  vals_.assign(v, v + nnz_);
}

template <typename T, typename I, typename A>
void CSRMatrix<T, I, A>::write_MatrixMarket(const std::string& fname) const {
  std::vector<int64_t> rp(row_ptr_.begin(), row_ptr_.end());
  std::vector<int32_t> col(col_ind_.begin(), col_ind_.end());
  bclx_csr c{(int64_t)m_, (int64_t)n_, (int64_t)nnz_, 8, detail::dtype_of<T>(), rp.data(), col.data(),
             const_cast<T*>(vals_.data())};
  detail::check(bclx_mm_write(fname.c_str(), &c));
}

// BCL's binary CSR file mapped read-only into memory (CSRMatrixMemoryMapped, CSRMatrixMemoryMapped.hpp:55-145): the
// arrays are views of the file, paged in as they are used; get_slice_impl_ copies one tile out.
template <typename T, typename I>
class CSRMatrixMemoryMapped {
 public:
  using value_type = T;
  using index_type = I;
  using size_type = size_t;

  explicit CSRMatrixMemoryMapped(const std::string& fname) {
    fd_ = open(fname.c_str(), O_RDONLY);
    if (fd_ < 0) throw std::runtime_error("CSRMatrixMemoryMapped: could not open file \"" + fname + "\"");
    struct stat st;
    if (fstat(fd_, &st) != 0 || (size_t)st.st_size < 3 * sizeof(size_t)) {
      close(fd_);
      throw std::runtime_error("CSRMatrixMemoryMapped: \"" + fname + "\" is not a binary CSR file");
    }
    length_ = (size_t)st.st_size;
    void* a = mmap(nullptr, length_, PROT_READ, MAP_SHARED, fd_, 0);
    if (a == MAP_FAILED) {
      close(fd_);
      throw std::runtime_error("CSRMatrixMemoryMapped: could not map \"" + fname + "\"");
    }
    address_ = a;
    const size_t* h = static_cast<const size_t*>(address_);
    m_ = h[0];
    n_ = h[1];
    nnz_ = h[2];
    if (3 * sizeof(size_t) + nnz_ * (sizeof(T) + sizeof(I)) + (m_ + 1) * sizeof(I) > length_) {
      munmap(address_, length_);
      close(fd_);
      throw std::runtime_error("CSRMatrixMemoryMapped: \"" + fname + "\" is shorter than its header says");
    }
    values_ = reinterpret_cast<const T*>(h + 3);
    colind_ = reinterpret_cast<const I*>(values_ + nnz_);
    rowptr_ = colind_ + nnz_;
  }
  CSRMatrixMemoryMapped(const CSRMatrixMemoryMapped&) = delete;
  CSRMatrixMemoryMapped& operator=(const CSRMatrixMemoryMapped&) = delete;
  ~CSRMatrixMemoryMapped() {
    munmap(address_, length_);
    close(fd_);
  }

  const T* values_data() const { return values_; }
  const I* colind_data() const { return colind_; }
  const I* rowptr_data() const { return rowptr_; }
  size_t m() const { return m_; }
  size_t n() const { return n_; }
  size_t nnz() const { return nnz_; }

  CSRMatrix<T, I> get_slice_impl_(size_type imin, size_type imax, size_type jmin, size_type jmax) const {
    imax = std::min(imax, m_);
    jmax = std::min(jmax, n_);
    CSRMatrix<T, I> r(imax - imin, jmax - jmin);
    for (size_type i = imin; i < imax; i++) {
      for (I p = rowptr_[i]; p < rowptr_[i + 1]; p++) {
        const size_type j = (size_type)colind_[p];
        if (j >= jmin && j < jmax) {
          r.vals_.push_back(values_[p]);
          r.col_ind_.push_back((I)(j - jmin));
        }
      }
      r.row_ptr_[i - imin + 1] = (I)r.vals_.size();
    }
    r.nnz_ = r.vals_.size();
    return r;
  }

 private:
  int fd_ = -1;
  void* address_ = nullptr;
  size_t length_ = 0;
  size_t m_ = 0, n_ = 0, nnz_ = 0;
  const T* values_ = nullptr;
  const I* colind_ = nullptr;
  const I* rowptr_ = nullptr;
};

// The tiles of a on a grid of tile_m x tile_n tiles, cut on the GPU (row-major over the grid): the slicing of
// SPMatrix's constructor (SPMatrix.hpp:212-266).
template <class M, typename = std::enable_if_t<detail::is_csr<M>::value>>
std::vector<M> split_grid(const M& a, size_t tile_m, size_t tile_n) {
  using T = typename M::value_type;
  using I = typename M::index_type;
  bclx_handle* h = detail::ctx().handle();
  detail::DevCSR<T, I> A(a);
  const size_t gm = std::max<size_t>(1, (a.m_ + tile_m - 1) / tile_m), gn = std::max<size_t>(1, (a.n_ + tile_n - 1) / tile_n);
  std::vector<bclx_csr> d(gm * gn);
  for (auto& x : d) x = bclx_csr{};
  detail::check(bclx_csr_split_grid(h, &A.d, (int64_t)tile_m, (int64_t)tile_n, (int32_t)d.size(), d.data()));
  std::vector<std::unique_ptr<detail::Dev<I>>> rp;
  std::vector<std::unique_ptr<detail::Dev<int32_t>>> col;
  std::vector<std::unique_ptr<detail::Dev<T>>> val;
  for (auto& x : d) {
    rp.push_back(std::make_unique<detail::Dev<I>>((size_t)x.m + 1));
    col.push_back(std::make_unique<detail::Dev<int32_t>>((size_t)x.nnz));
    val.push_back(std::make_unique<detail::Dev<T>>((size_t)x.nnz));
    x.ptr_width = detail::width_of<I>();
    x.row_ptr = rp.back()->p;
    x.col_idx = col.back()->p;
    x.vals = val.back()->p;
  }
  detail::check(bclx_csr_split_grid(h, &A.d, (int64_t)tile_m, (int64_t)tile_n, (int32_t)d.size(), d.data()));
  detail::check(bclx_synchronize(h));
  std::vector<M> out;
  for (size_t t = 0; t < d.size(); t++)
    out.push_back(detail::to_host<M>(*rp[t], *col[t], *val[t], (size_t)d[t].m, (size_t)d[t].n, (size_t)d[t].nnz));
  return out;
}

// Elementwise tile arithmetic of DMatrix (DMatrix.hpp:319-386) on host tiles, run on the GPU: c = a op b for
// op in {'+', '-', '*', '/'}; c may be a (binary_op_inplace, operator+=).
template <typename T>
void tile_binary_op(char op, size_t m, size_t n, const T* a, size_t lda, const T* b, size_t ldb, T* c, size_t ldc) {
  const int code = op == '+' ? 0 : op == '-' ? 1 : op == '*' ? 2 : op == '/' ? 3 : -1;
  if (code < 0) throw std::runtime_error("bclx: tile_binary_op takes '+', '-', '*' or '/'");
  bclx_handle* h = detail::ctx().handle();
  detail::Dev<T> A(a, m * lda), B(b, m * ldb), C(m * ldc);
  if (c != a && ldc > n) detail::hip_check(hipMemcpy(C.p, c, sizeof(T) * m * ldc, hipMemcpyHostToDevice), "hipMemcpy");
  detail::check(bclx_dense_binary_op(h, code, detail::dtype_of<T>(), (int64_t)m, (int64_t)n, A.p, (int64_t)lda, B.p,
                                     (int64_t)ldb, c == a ? A.p : C.p, (int64_t)(c == a ? lda : ldc)));
  detail::check(bclx_synchronize(h));
  (c == a ? A : C).to_host(c, m * (c == a ? lda : ldc));
}

// c += a * b (BCL::gemm(CSRMatrix, CSRMatrix, CSRMatrix)): the existing entries of c are kept and summed with the
// product; new entries are created for every product position, including exact-zero sums. M is bclx::CSRMatrix or
// BCL::CSRMatrix<T, I, Allocator> itself.
template <class M, typename = std::enable_if_t<detail::is_csr<M>::value>>
void gemm(const M& a, const M& b, M& c) {
  if (c.m_ != a.m_ || c.n_ != b.n_) throw std::runtime_error("bclx: c has the wrong shape");
  M p = detail::multiply(a, b);
  if (c.nnz_ == 0) {
    c = std::move(p);
    return;
  }
  c = detail::sum<M>({&c, &p}, c.m_, c.n_);
}

// BCL::SparseHashAccumulator<T, I, Allocator> (SparseHashAccumulator.hpp:86-183): collect partial C tiles, then
// get_matrix(m, n) sums them on the GPU. M is the tile type (BCL::CSRMatrix<T, I, Allocator> works as well).
template <typename T, typename I = int, typename Allocator = std::allocator<T>, class M = CSRMatrix<T, I, Allocator>>
struct SparseHashAccumulator {
  std::vector<M> mats_;
  void accumulate(M&& mat) { mats_.push_back(std::move(mat)); }
  M get_matrix(size_t m, size_t n) const {
    if (mats_.empty()) {
      decltype(M::vals_) v;
      decltype(M::row_ptr_) r(m + 1, 0);
      decltype(M::col_ind_) c;
      return M(m, n, 0, std::move(v), std::move(r), std::move(c));
    }
    std::vector<const M*> p;
    for (auto& x : mats_) p.push_back(&x);
    return detail::sum<M>(p, m, n);
  }
};

// BCL::spmm_wrapper (spgemm.hpp:140-149): c[i*ldc + j] += sum_k a[i,k] * b[k*ldb + j], j < n; host arrays, T = float
// (fp32 accumulation) or double (fp64 throughout, the reference's double path, spgemm.hpp:119-138).
template <typename I, typename T>
void spmm_wrapper(size_t m, size_t n, size_t k, size_t nnz, T* c, T* b, T* a_values, I* a_rowptr, I* a_colind,
                  size_t ldb, size_t ldc) {
  bclx_handle* h = detail::ctx().handle();
  detail::Dev<I> rp(a_rowptr, m + 1);
  std::vector<int32_t> ci(a_colind, a_colind + nnz);
  detail::Dev<int32_t> col(ci.data(), nnz);
  detail::Dev<T> val(a_values, nnz);
  detail::Dev<T> B(b, k * ldb), C(c, m * ldc);
  bclx_csr A{(int64_t)m, (int64_t)k, (int64_t)nnz, detail::width_of<I>(), detail::dtype_of<T>(), rp.p, col.p, val.p};
  if constexpr (std::is_same<T, double>::value)
    detail::check(bclx_spmm_f64(h, &A, B.p, (int64_t)ldb, C.p, (int64_t)ldc, (int64_t)n, 1.0, 1.0));
  else
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

// BCL::experimental::cblas_gemm_wrapper_ (double -> cblas_dgemm, cblas_wrapper.hpp:29-40): fp64 MFMA.
inline void cblas_gemm_wrapper_(int layout, int transa, int transb, int m, int n, int k, double alpha, const double* a,
                                int lda, const double* b, int ldb, double beta, double* c, int ldc) {
  bclx_handle* h = detail::ctx().handle();
  const bool ta = transa != NoTrans, tb = transb != NoTrans;
  if (layout == ColMajor) {
    std::swap(m, n);
    std::swap(a, b);
    std::swap(lda, ldb);
    const bool t = ta;
    return cblas_gemm_wrapper_(RowMajor, tb ? Trans : NoTrans, t ? Trans : NoTrans, m, n, k, alpha, a, lda, b, ldb,
                               beta, c, ldc);
  }
  const size_t a_rows = ta ? k : m, b_rows = tb ? n : k;
  detail::Dev<double> A(a, a_rows * (size_t)lda), B(b, b_rows * (size_t)ldb), C(c, (size_t)m * ldc);
  detail::check(bclx_gemm_f64(h, ta, tb, m, n, k, alpha, A.p, lda, B.p, ldb, beta, C.p, ldc));
  detail::check(bclx_synchronize(h));
  C.to_host(c, (size_t)m * ldc);
}

}  // namespace bclx
