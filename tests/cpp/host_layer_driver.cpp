// Drives the C++ host layer (include/bclx/bclx.hpp) the way BCL code calls the local tile multiply, on files
// written by tests/test_cpp_layer_gpu.py in BCL's binary CSR format:
//   <dir>/a.binary, b.binary, c0.binary -> dot.bin = a.dot(b) (also dot_alloc.bin with a custom allocator, and
//                                   dot_like.bin / gemm_like.bin / acc_like.bin on a type with only BCL's members);  gemm.bin = c0 after bclx::gemm(a, b, c0)  (c0 += a * b)
//                                   acc.bin = SparseHashAccumulator over {a*b, c0}
//   <dir>/dense.bin (m, n, k, then A m x k, B k x n, C m x n as float) -> gemm_out.bin = C + A * B (cblas wrapper,
//   row-major), gemm_cm.bin = the same product through the column-major path
//   spmm: spmm_in.bin (n, ld, then B k x ld, C m x ld) with a -> spmm_out.bin (C += a * B[:, :n])
//   fp64: dense64.bin (m, n, k, then A, B, C as double) -> gemm64_out.bin (row-major result, then column-major);
//         spmm64_out.bin = C + a * B in fp64 (a's values, B and C of spmm_in.bin widened)
//   files: a.mtx -> mtx.binary (CSRMatrix(fname) of a MatrixMarket file), mtx_out.mtx (write_MatrixMarket);
//   a.binary memory-mapped -> slice.binary = rows [100, 700) x columns [50, 1500); split_grid(a, 300, 500) ->
//   split_<t>.binary; ew.bin (m, n, ld, then X, Y m x ld) -> ew_out.bin = X + Y, X - Y, X * Y, X / Y, then X += Y
#include <bclx/bclx.hpp>

#include <cstdio>
#include <string>
#include <vector>

// Containers other than bclx::CSRMatrix<float, int>: one with a custom allocator (the reference's Allocator template
// parameter), and a type that only has BCL::CSRMatrix's public members and constructor, as BCL's own tiles do.
template <class T>
struct TrackingAlloc {
  using value_type = T;
  static inline size_t live = 0;
  TrackingAlloc() = default;
  template <class U>
  TrackingAlloc(const TrackingAlloc<U>&) {}
  T* allocate(size_t n) {
    live += n;
    return std::allocator<T>().allocate(n);
  }
  void deallocate(T* p, size_t n) {
    live -= n;
    std::allocator<T>().deallocate(p, n);
  }
  template <class U>
  bool operator==(const TrackingAlloc<U>&) const { return true; }
  template <class U>
  bool operator!=(const TrackingAlloc<U>&) const { return false; }
};

struct LikeBCL {
  using value_type = float;
  using index_type = int;
  size_t m_, n_, nnz_;
  std::vector<float> vals_;
  std::vector<int> row_ptr_, col_ind_;
  LikeBCL(size_t m, size_t n, size_t nnz, std::vector<float>&& v, std::vector<int>&& r, std::vector<int>&& c)
      : m_(m), n_(n), nnz_(nnz), vals_(std::move(v)), row_ptr_(std::move(r)), col_ind_(std::move(c)) {}
};

template <class M, class S>
static M convert(const S& x) {
  decltype(M::vals_) v(x.vals_.begin(), x.vals_.end());
  decltype(M::row_ptr_) r(x.row_ptr_.begin(), x.row_ptr_.end());
  decltype(M::col_ind_) c(x.col_ind_.begin(), x.col_ind_.end());
  return M(x.m_, x.n_, x.nnz_, std::move(v), std::move(r), std::move(c));
}

template <typename T>
static std::vector<T> read_vec(FILE* f, size_t n) {
  std::vector<T> v(n);
  if (n && fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short read");
  return v;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s DIR\n", argv[0]);
    return 2;
  }
  const std::string d = argv[1];
  try {
    bclx::CSRMatrix<float, int> a(d + "/a.binary"), b(d + "/b.binary"), c0(d + "/c0.binary");
    a.dot(b).write_Binary(d + "/dot.bin");
    bclx::CSRMatrix<float, int> c = c0;
    bclx::gemm(a, b, c);
    c.write_Binary(d + "/gemm.bin");
    bclx::SparseHashAccumulator<float, int> acc;
    acc.accumulate(a.dot(b));
    acc.accumulate(bclx::CSRMatrix<float, int>(c0));
    acc.get_matrix(a.m(), b.n()).write_Binary(d + "/acc.bin");
    {  // the same products on other container types
      using AllocM = bclx::CSRMatrix<float, int, TrackingAlloc<float>>;
      AllocM ta = convert<AllocM>(a), tb = convert<AllocM>(b);
      convert<bclx::CSRMatrix<float, int>>(ta.dot(tb)).write_Binary(d + "/dot_alloc.bin");
      if (TrackingAlloc<float>::live == 0) throw std::runtime_error("the custom allocator was not used");
      LikeBCL la = convert<LikeBCL>(a), lb = convert<LikeBCL>(b), lc = convert<LikeBCL>(c0);
      convert<bclx::CSRMatrix<float, int>>(bclx::dot(la, lb)).write_Binary(d + "/dot_like.bin");
      bclx::gemm(la, lb, lc);
      convert<bclx::CSRMatrix<float, int>>(lc).write_Binary(d + "/gemm_like.bin");
      bclx::SparseHashAccumulator<float, int, std::allocator<float>, LikeBCL> lacc;
      lacc.accumulate(bclx::dot(la, lb));
      lacc.accumulate(convert<LikeBCL>(c0));
      convert<bclx::CSRMatrix<float, int>>(lacc.get_matrix(a.m(), b.n())).write_Binary(d + "/acc_like.bin");
    }

    FILE* f = fopen((d + "/dense.bin").c_str(), "rb");
    auto dims = read_vec<int64_t>(f, 3);
    const int m = (int)dims[0], n = (int)dims[1], k = (int)dims[2];
    auto A = read_vec<float>(f, (size_t)m * k), B = read_vec<float>(f, (size_t)k * n), C = read_vec<float>(f, (size_t)m * n);
    fclose(f);
    std::vector<float> Cr = C, Cc(C.size());
    bclx::cblas_gemm_wrapper_(bclx::RowMajor, bclx::NoTrans, bclx::NoTrans, m, n, k, 1.f, A.data(), k, B.data(), n,
                              1.f, Cr.data(), n);
    // column-major: the same product on transposed copies, result transposed back
    std::vector<float> At((size_t)k * m), Bt((size_t)n * k), Ct((size_t)n * m);
    for (int i = 0; i < m; i++)
      for (int q = 0; q < k; q++) At[(size_t)q * m + i] = A[(size_t)i * k + q];
    for (int q = 0; q < k; q++)
      for (int j = 0; j < n; j++) Bt[(size_t)j * k + q] = B[(size_t)q * n + j];
    for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) Ct[(size_t)j * m + i] = C[(size_t)i * n + j];
    bclx::cblas_gemm_wrapper_(bclx::ColMajor, bclx::NoTrans, bclx::NoTrans, m, n, k, 1.f, At.data(), m, Bt.data(), k,
                              1.f, Ct.data(), m);
    for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) Cc[(size_t)i * n + j] = Ct[(size_t)j * m + i];
    FILE* o = fopen((d + "/gemm_out.bin").c_str(), "wb");
    fwrite(Cr.data(), sizeof(float), Cr.size(), o);
    fclose(o);
    o = fopen((d + "/gemm_cm.bin").c_str(), "wb");
    fwrite(Cc.data(), sizeof(float), Cc.size(), o);
    fclose(o);

    f = fopen((d + "/spmm_in.bin").c_str(), "rb");
    auto sd = read_vec<int64_t>(f, 2);
    const size_t sn = (size_t)sd[0], ld = (size_t)sd[1];
    auto SB = read_vec<float>(f, a.n() * ld), SC = read_vec<float>(f, a.m() * ld);
    fclose(f);
    bclx::spmm_wrapper<int>(a.m(), sn, a.n(), a.nnz(), SC.data(), SB.data(), a.vals_.data(), a.row_ptr_.data(),
                            a.col_ind_.data(), ld, ld);
    o = fopen((d + "/spmm_out.bin").c_str(), "wb");
    fwrite(SC.data(), sizeof(float), SC.size(), o);
    fclose(o);

    {  // the double overloads: cblas_gemm_wrapper_ (row- and column-major) and spmm_wrapper<int, double>
      f = fopen((d + "/dense64.bin").c_str(), "rb");
      auto dd = read_vec<int64_t>(f, 3);
      const int m6 = (int)dd[0], n6 = (int)dd[1], k6 = (int)dd[2];
      auto A6 = read_vec<double>(f, (size_t)m6 * k6), B6 = read_vec<double>(f, (size_t)k6 * n6);
      auto C6 = read_vec<double>(f, (size_t)m6 * n6);
      fclose(f);
      std::vector<double> R6 = C6;
      bclx::cblas_gemm_wrapper_(bclx::RowMajor, bclx::NoTrans, bclx::NoTrans, m6, n6, k6, 1.0, A6.data(), k6,
                                B6.data(), n6, 1.0, R6.data(), n6);
      std::vector<double> At6((size_t)k6 * m6), Bt6((size_t)n6 * k6), Ct6((size_t)n6 * m6), M6(C6.size());
      for (int i = 0; i < m6; i++)
        for (int q = 0; q < k6; q++) At6[(size_t)q * m6 + i] = A6[(size_t)i * k6 + q];
      for (int q = 0; q < k6; q++)
        for (int j = 0; j < n6; j++) Bt6[(size_t)j * k6 + q] = B6[(size_t)q * n6 + j];
      for (int i = 0; i < m6; i++)
        for (int j = 0; j < n6; j++) Ct6[(size_t)j * m6 + i] = C6[(size_t)i * n6 + j];
      bclx::cblas_gemm_wrapper_(bclx::ColMajor, bclx::NoTrans, bclx::NoTrans, m6, n6, k6, 1.0, At6.data(), m6,
                                Bt6.data(), k6, 1.0, Ct6.data(), m6);
      for (int i = 0; i < m6; i++)
        for (int j = 0; j < n6; j++) M6[(size_t)i * n6 + j] = Ct6[(size_t)j * m6 + i];
      o = fopen((d + "/gemm64_out.bin").c_str(), "wb");
      fwrite(R6.data(), sizeof(double), R6.size(), o);
      fwrite(M6.data(), sizeof(double), M6.size(), o);
      fclose(o);
      // spmm in fp64 on a with its values widened; B and C from spmm_in.bin widened
      std::vector<double> av(a.vals_.begin(), a.vals_.end());
      std::vector<double> B64(SB.begin(), SB.end()), C64(SC.begin(), SC.end());
      f = fopen((d + "/spmm_in.bin").c_str(), "rb");  // the original C (SC now holds the fp32 result)
      read_vec<int64_t>(f, 2);
      read_vec<float>(f, a.n() * ld);
      auto SC0 = read_vec<float>(f, a.m() * ld);
      fclose(f);
      std::copy(SC0.begin(), SC0.end(), C64.begin());
      bclx::spmm_wrapper<int>(a.m(), sn, a.n(), a.nnz(), C64.data(), B64.data(), av.data(), a.row_ptr_.data(),
                              a.col_ind_.data(), ld, ld);
      o = fopen((d + "/spmm64_out.bin").c_str(), "wb");
      fwrite(C64.data(), sizeof(double), C64.size(), o);
      fclose(o);
    }

    bclx::CSRMatrix<float, int> mtx(d + "/a.mtx");
    mtx.write_Binary(d + "/mtx.binary");
    mtx.write_MatrixMarket(d + "/mtx_out.mtx");
    auto info = bclx::matrix_io::matrix_info(d + "/a.mtx");
    if (info.shape[0] != mtx.m() || info.shape[1] != mtx.n()) throw std::runtime_error("matrix_info disagrees");
    bclx::CSRMatrixMemoryMapped<float, int> mm(d + "/a.binary");
    mm.get_slice_impl_(100, 700, 50, 1500).write_Binary(d + "/slice.binary");
    auto tiles = bclx::split_grid(a, 300, 500);
    for (size_t t = 0; t < tiles.size(); t++) tiles[t].write_Binary(d + "/split_" + std::to_string(t) + ".binary");

    f = fopen((d + "/ew.bin").c_str(), "rb");
    auto ed = read_vec<int64_t>(f, 3);
    const size_t em = (size_t)ed[0], en = (size_t)ed[1], eld = (size_t)ed[2];
    auto X = read_vec<double>(f, em * eld), Y = read_vec<double>(f, em * eld);
    fclose(f);
    o = fopen((d + "/ew_out.bin").c_str(), "wb");
    for (char op : std::string("+-*/")) {
      std::vector<double> Z(em * eld, -1.0);
      bclx::tile_binary_op<double>(op, em, en, X.data(), eld, Y.data(), eld, Z.data(), eld);
      fwrite(Z.data(), sizeof(double), Z.size(), o);
    }
    bclx::tile_binary_op<double>('+', em, en, X.data(), eld, Y.data(), eld, X.data(), eld);  // X += Y
    fwrite(X.data(), sizeof(double), X.size(), o);
    fclose(o);
  } catch (const std::exception& e) {
    fprintf(stderr, "host_layer_driver: %s\n", e.what());
    return 1;
  }
  printf("host_layer_driver ok\n");
  return 0;
}
