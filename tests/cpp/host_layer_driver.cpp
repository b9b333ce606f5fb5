// Drives the C++ host layer (include/bclx/bclx.hpp) the way BCL code calls the local tile multiply, on files
// written by tests/test_cpp_layer_gpu.py in BCL's binary CSR format:
//   <dir>/a.bin, b.bin, c0.bin  ->  dot.bin = a.dot(b);  gemm.bin = c0 after bclx::gemm(a, b, c0)  (c0 += a * b)
//                                   acc.bin = SparseHashAccumulator over {a*b, c0}
//   <dir>/dense.bin (m, n, k, then A m x k, B k x n, C m x n as float) -> gemm_out.bin = C + A * B (cblas wrapper,
//   row-major), gemm_cm.bin = the same product through the column-major path
//   spmm: spmm_in.bin (n, ld, then B k x ld, C m x ld) with a -> spmm_out.bin (C += a * B[:, :n])
#include <bclx/bclx.hpp>

#include <cstdio>
#include <string>
#include <vector>

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
    bclx::CSRMatrix<float, int> a(d + "/a.bin"), b(d + "/b.bin"), c0(d + "/c0.bin");
    a.dot(b).write_Binary(d + "/dot.bin");
    bclx::CSRMatrix<float, int> c = c0;
    bclx::gemm(a, b, c);
    c.write_Binary(d + "/gemm.bin");
    bclx::SparseHashAccumulator<float, int> acc;
    acc.accumulate(a.dot(b));
    acc.accumulate(bclx::CSRMatrix<float, int>(c0));
    acc.get_matrix(a.m(), b.n()).write_Binary(d + "/acc.bin");

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
  } catch (const std::exception& e) {
    fprintf(stderr, "host_layer_driver: %s\n", e.what());
    return 1;
  }
  printf("host_layer_driver ok\n");
  return 0;
}
