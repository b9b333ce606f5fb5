// SpMM: C[i*ldc + j] = beta * C + alpha * sum_k A[i,k] * B[k*ldb + j], j < n (bclx_spmm).
//
// The reference's spmm_wrapper (bcl/containers/algorithms/spgemm.hpp:78-96, :140-149) is alpha = beta = 1 with
// row-major B and C and explicit ldb / ldc (DMatrix tiles padded to the full tile width).
//
// Work split: A's nonzeros are cut into chunks of kChunk consecutive entries; one wavefront takes a chunk, finds the
// row holding its first entry by binary search in row_ptr, and walks the chunk's entries in order. Each lane owns
// VEC consecutive columns of a 64*VEC column pass; for every entry the wavefront gathers one B row segment
// (coalesced; the loads of eight entries are issued together), multiplies by the entry's value (broadcast from the
// lane that loaded it) and accumulates in fp32.
// Each finished row (or a chunk's part of a row cut by a chunk boundary) is added to C with atomics that return
// nothing, so no wavefront waits on a read of C. B is fp32 or bf16 (converted on load) with fp32 accumulation, or,
// for the fp64 product (bclx_spmm_f64, the reference's double spmm_wrapper, spgemm.hpp:119-149), fp64 throughout.
// bf16 B takes a variant in which one 8-byte load per lane serves two entries (k_spmm_bf16_pair).
#include <algorithm>

#include "internal.hpp"

namespace bclx {

constexpr int kSpmmChunk = 256;  // entries per wavefront
constexpr int kSpmmWaves = 4;    // wavefronts per block
constexpr int kSpmmUnroll = 8;   // B-row loads in flight per wavefront

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <typename BT, int VEC>
struct BLoad;
template <>
struct BLoad<double, 2> {
  static __device__ inline void load(const double* p, double* o) {
    const double2 v = *reinterpret_cast<const double2*>(p);
    o[0] = v.x;
    o[1] = v.y;
  }
};
template <>
struct BLoad<float, 2> {
  static __device__ inline void load(const float* p, float* o) {
    const float2 v = *reinterpret_cast<const float2*>(p);
    o[0] = v.x;
    o[1] = v.y;
  }
};
template <>
struct BLoad<uint16_t, 2> {
  static __device__ inline void load(const uint16_t* p, float* o) {
    const uint32_t v = *reinterpret_cast<const uint32_t*>(p);
    o[0] = __uint_as_float(v << 16);
    o[1] = __uint_as_float(v & 0xFFFF0000u);
  }
};
template <>
struct BLoad<uint16_t, 4> {
  static __device__ inline void load(const uint16_t* p, float* o) {
    const u32x2 v = *reinterpret_cast<const u32x2*>(p);
    o[0] = __uint_as_float(v[0] << 16);
    o[1] = __uint_as_float(v[0] & 0xFFFF0000u);
    o[2] = __uint_as_float(v[1] << 16);
    o[3] = __uint_as_float(v[1] & 0xFFFF0000u);
  }
};
template <typename BT>
struct BLoad<BT, 1> {
  template <typename VT>
  static __device__ inline void load(const BT* p, VT* o) {
    if constexpr (sizeof(BT) == 2) o[0] = bf16_to_f32(*p);
    else o[0] = *p;
  }
};

// C segment += alpha * acc with fp32 atomics that return nothing: the wavefront never waits on a read of C, and a row
// split between chunks sums its parts in L2.
template <int VEC, typename VT>
__device__ inline void spmm_flush(VT* cr, const VT* acc, VT alpha, int left) {
#pragma unroll
  for (int v = 0; v < VEC; v++)
    if (v < left) __hip_atomic_fetch_add(&cr[v], alpha * acc[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// VT: A's values, the accumulation and C (float, or double for the fp64 product); BT: B (VT, or bf16 under float).
template <typename RP, typename VT, typename BT, int VEC>
__global__ __launch_bounds__(64 * kSpmmWaves) void k_spmm(const RP* __restrict__ rp, const int* __restrict__ col,
                                                          const VT* __restrict__ val, int64_t m, int64_t nnz,
                                                          const BT* __restrict__ B, int64_t ldb, VT* C, int64_t ldc,
                                                          int n, VT alpha) {
  const int lane = threadIdx.x & 63;
  const int64_t chunk = (int64_t)blockIdx.x * kSpmmWaves + (threadIdx.x >> 6);
  const int64_t e0 = chunk * kSpmmChunk;
  if (e0 >= nnz) return;
  const int64_t e1 = e0 + kSpmmChunk < nnz ? e0 + kSpmmChunk : nnz;
  // first row with rp[r + 1] > e0
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)rp[mid + 1] <= e0) lo = mid + 1;
    else hi = mid;
  }
  // row ends of the chunk's first 64 rows, one per lane (rows further on are read when reached)
  const int64_t my_re = lo + 1 + lane <= m ? (int64_t)rp[lo + 1 + lane] : nnz;
  for (int c0 = 0; c0 < n; c0 += 64 * VEC) {
    const int cj = c0 + lane * VEC;
    const bool cok = cj < n;
    int64_t r = lo;
    int64_t rs = (int64_t)rp[r], re = __shfl(my_re, 0, 64);
    VT acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; v++) acc[v] = VT(0);
    for (int64_t eb = e0; eb < e1; eb += 64) {
      // stage 64 entries: lane q holds entry eb + q
      const int64_t eq = eb + lane;
      const int my_col = eq < e1 ? col[eq] : 0;
      const VT my_val = eq < e1 ? val[eq] : VT(0);
      const int cnt = (int)((e1 - eb) < 64 ? (e1 - eb) : 64);
      for (int q0 = 0; q0 < cnt; q0 += kSpmmUnroll) {
        // issue the B-row loads of kSpmmUnroll entries before using any, so their latencies overlap
        VT b[kSpmmUnroll][VEC];
        VT a[kSpmmUnroll];
#pragma unroll
        for (int u = 0; u < kSpmmUnroll; u++) {
          const int q = q0 + u < cnt ? q0 + u : cnt - 1;
          const int k = __shfl(my_col, q, 64);
          a[u] = __shfl(my_val, q, 64);
#pragma unroll
          for (int v = 0; v < VEC; v++) b[u][v] = VT(0);
          if (cok && q0 + u < cnt) {
            if (cj + VEC <= n) {
              BLoad<BT, VEC>::load(B + (int64_t)k * ldb + cj, b[u]);
            } else {
              for (int v = 0; v < VEC && cj + v < n; v++) BLoad<BT, 1>::load(B + (int64_t)k * ldb + cj + v, &b[u][v]);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kSpmmUnroll; u++) {
          if (q0 + u >= cnt) break;
          const int64_t e = eb + q0 + u;
          while (e >= re) {  // flush the finished row (empty rows have nothing to add), move to the row holding e
            if (cok && re > rs) spmm_flush<VEC>(C + r * ldc + cj, acc, alpha, n - cj);
#pragma unroll
            for (int v = 0; v < VEC; v++) acc[v] = VT(0);
            r++;
            rs = re;
            re = r - lo < 64 ? __shfl(my_re, (int)(r - lo), 64) : (int64_t)rp[r + 1];
          }
#pragma unroll
          for (int v = 0; v < VEC; v++) acc[v] = fma(a[u], b[u][v], acc[v]);
        }
      }
    }
    if (cok) spmm_flush<VEC>(C + r * ldc + cj, acc, alpha, n - cj);  // the chunk's last (possibly partial) row
  }
}

// bf16 B with paired entries: each lane loads 8 bytes (four bf16 columns), so half a wavefront covers a 128-column
// pass of one B row and one load instruction serves two consecutive entries: lanes 0-31 take the even entry of a
// pair, lanes 32-63 the odd one. (With 4-byte loads the bf16 kernel issues as many loads as the fp32 one for half the
// bytes.) Both halves accumulate the same row; when the odd entry starts a new row, the even half's product is added
// before the row is flushed. A flush sums the two halves (lanes l and l + 32 hold the same columns), and each half
// adds two of its four columns to C with atomics. Needs n and ldb multiples of 4 and B aligned to 8 bytes.
template <typename RP, int U>
__global__ __launch_bounds__(64 * kSpmmWaves) void k_spmm_bf16_pair(const RP* __restrict__ rp,
                                                                    const int* __restrict__ col,
                                                                    const float* __restrict__ val, int64_t m,
                                                                    int64_t nnz, const uint16_t* __restrict__ B,
                                                                    int64_t ldb, float* C, int64_t ldc, int n,
                                                                    float alpha) {
  constexpr int VEC = 4;
  const int lane = threadIdx.x & 63, half = lane >> 5, hl = lane & 31;
  const int64_t chunk = (int64_t)blockIdx.x * kSpmmWaves + (threadIdx.x >> 6);
  const int64_t e0 = chunk * kSpmmChunk;
  if (e0 >= nnz) return;
  const int64_t e1 = e0 + kSpmmChunk < nnz ? e0 + kSpmmChunk : nnz;
  int64_t lo = 0, hi = m;  // first row with rp[r + 1] > e0
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)rp[mid + 1] <= e0) lo = mid + 1;
    else hi = mid;
  }
  const int64_t my_re = lo + 1 + lane <= m ? (int64_t)rp[lo + 1 + lane] : nnz;
  for (int c0 = 0; c0 < n; c0 += 32 * VEC) {
    const int cj = c0 + hl * VEC;
    const bool cok = cj < n;  // n is a multiple of 4: a lane's columns are all in range or all out
    int64_t r = lo;
    int64_t rs = (int64_t)rp[r], re = __shfl(my_re, 0, 64);
    float acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; v++) acc[v] = 0.f;
    // add row r's sum (both halves) to C; every lane runs the exchange, then half h adds columns 2h, 2h + 1
    auto add_row = [&](bool nonempty) {
      float tot[VEC / 2];
#pragma unroll
      for (int v = 0; v < VEC / 2; v++) {
        const float mine = half ? acc[VEC / 2 + v] : acc[v];
        const float other = __shfl_xor(half ? acc[v] : acc[VEC / 2 + v], 32, 64);
        tot[v] = mine + other;
      }
      if (cok && nonempty) {
        float* cr = C + r * ldc + cj + half * (VEC / 2);
#pragma unroll
        for (int v = 0; v < VEC / 2; v++)
          __hip_atomic_fetch_add(&cr[v], alpha * tot[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    };
    auto next_row = [&]() {
      add_row(re > rs);
#pragma unroll
      for (int v = 0; v < VEC; v++) acc[v] = 0.f;
      r++;
      rs = re;
      const int64_t sh = __shfl(my_re, (int)(r - lo < 64 ? r - lo : 63), 64);
      re = r - lo < 64 ? sh : (int64_t)rp[r + 1];
    };
    for (int64_t eb = e0; eb < e1; eb += 64) {
      const int64_t eq = eb + lane;
      const int my_col = eq < e1 ? col[eq] : 0;
      const float my_val = eq < e1 ? val[eq] : 0.f;
      const int cnt = (int)((e1 - eb) < 64 ? (e1 - eb) : 64);
      for (int p0 = 0; p0 < cnt; p0 += 2 * U) {
        float b[U][VEC];
        float a[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int q = p0 + 2 * u + half;
          const int k = __shfl(my_col, q < cnt ? q : cnt - 1, 64);
          a[u] = __shfl(my_val, q < cnt ? q : cnt - 1, 64);
#pragma unroll
          for (int v = 0; v < VEC; v++) b[u][v] = 0.f;
          if (cok && q < cnt) BLoad<uint16_t, VEC>::load(B + (int64_t)k * ldb + cj, b[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int q = p0 + 2 * u;  // the pair's even entry (wave-uniform)
          if (q >= cnt) break;
          const int64_t e = eb + q;
          while (e >= re) next_row();
          const bool odd = q + 1 < cnt;
          const bool same = odd && e + 1 < re;
          if (half == 0 || same) {
#pragma unroll
            for (int v = 0; v < VEC; v++) acc[v] = fmaf(a[u], b[u][v], acc[v]);
          }
          if (odd && !same) {
            while (e + 1 >= re) next_row();
            if (half == 1) {
#pragma unroll
              for (int v = 0; v < VEC; v++) acc[v] = fmaf(a[u], b[u][v], acc[v]);
            }
          }
        }
      }
    }
    add_row(true);  // the chunk's last (possibly partial) row
  }
}

template <typename VT>
__global__ void k_scale_rows(VT* C, int64_t m, int64_t n, int64_t ldc, VT beta) {
  const int64_t total = m * n;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / n, j = t % n;
    VT* p = C + i * ldc + j;
    *p = beta == VT(0) ? VT(0) : beta * *p;
  }
}

template <typename VT>
static void scale_rows(bclx_handle* h, VT* C, int64_t m, int64_t n, int64_t ldc, VT beta) {
  if (beta == VT(1) || m == 0 || n == 0) return;
  const int64_t total = m * n;
  hipLaunchKernelGGL((k_scale_rows<VT>), dim3((unsigned)std::min<int64_t>((total + 255) / 256, 65536)), dim3(256), 0,
                     h->stream, C, m, n, ldc, beta);
  BCLX_HIP(hipGetLastError());
}

template <typename RP, typename VT, typename BT>
static void spmm_impl(bclx_handle* h, const bclx_csr* A, const void* Bv, int64_t ldb, VT* C, int64_t ldc, int64_t n,
                      VT alpha) {
  const RP* rp = static_cast<const RP*>(A->row_ptr);
  const BT* B = static_cast<const BT*>(Bv);
  const int64_t chunks = (A->nnz + kSpmmChunk - 1) / kSpmmChunk;
  if (chunks == 0 || n == 0) return;
  const unsigned grid = div_up(chunks, kSpmmWaves);
  const bool vec2 = (n % 2 == 0) && (ldb % 2 == 0) && (ldc % 2 == 0) &&
                    ((reinterpret_cast<uintptr_t>(B) % (2 * sizeof(BT))) == 0);
  const bool pair = sizeof(BT) == 2 && n % 4 == 0 && ldb % 4 == 0 && reinterpret_cast<uintptr_t>(B) % 8 == 0;
  if (pair) {
    if constexpr (sizeof(BT) == 2)
      hipLaunchKernelGGL((k_spmm_bf16_pair<RP, 4>), dim3(grid), dim3(64 * kSpmmWaves), 0, h->stream, rp, A->col_idx,
                         static_cast<const float*>(A->vals), A->m, A->nnz, B, ldb, C, ldc, (int)n, alpha);
  } else if (vec2)
    hipLaunchKernelGGL((k_spmm<RP, VT, BT, 2>), dim3(grid), dim3(64 * kSpmmWaves), 0, h->stream, rp, A->col_idx,
                       static_cast<const VT*>(A->vals), A->m, A->nnz, B, ldb, C, ldc, (int)n, alpha);
  else
    hipLaunchKernelGGL((k_spmm<RP, VT, BT, 1>), dim3(grid), dim3(64 * kSpmmWaves), 0, h->stream, rp, A->col_idx,
                       static_cast<const VT*>(A->vals), A->m, A->nnz, B, ldb, C, ldc, (int)n, alpha);
  BCLX_HIP(hipGetLastError());
}

}  // namespace bclx

using namespace bclx;

extern "C" {

bclx_status bclx_spmm(bclx_handle* h, const bclx_csr* A, const void* d_B, int32_t b_dtype, int64_t ldb, float* d_C,
                      int64_t ldc, int64_t n, float alpha, float beta) {
  return guarded([&] {
    BCLX_REQUIRE(h && A && d_C, BCLX_ERR_INVALID, "null argument");
    BCLX_REQUIRE(A->dtype == BCLX_F32, BCLX_ERR_INVALID, "A must hold fp32 values");
    BCLX_REQUIRE(A->ptr_width == 4 || A->ptr_width == 8, BCLX_ERR_INVALID, "A->ptr_width must be 4 or 8");
    BCLX_REQUIRE(b_dtype == BCLX_F32 || b_dtype == BCLX_BF16, BCLX_ERR_INVALID, "B must be fp32 or bf16");
    BCLX_REQUIRE(n >= 0 && ldb >= n && ldc >= n, BCLX_ERR_INVALID, "need ldb >= n and ldc >= n");
    BCLX_REQUIRE(A->nnz == 0 || (A->col_idx && A->vals && d_B), BCLX_ERR_INVALID, "null A arrays or B");
    set_device(h);
    scale_rows<float>(h, d_C, A->m, n, ldc, beta);
    if (A->ptr_width == 8) {
      if (b_dtype == BCLX_F32) spmm_impl<int64_t, float, float>(h, A, d_B, ldb, d_C, ldc, n, alpha);
      else spmm_impl<int64_t, float, uint16_t>(h, A, d_B, ldb, d_C, ldc, n, alpha);
    } else {
      if (b_dtype == BCLX_F32) spmm_impl<int32_t, float, float>(h, A, d_B, ldb, d_C, ldc, n, alpha);
      else spmm_impl<int32_t, float, uint16_t>(h, A, d_B, ldb, d_C, ldc, n, alpha);
    }
  });
}

bclx_status bclx_spmm_f64(bclx_handle* h, const bclx_csr* A, const double* d_B, int64_t ldb, double* d_C, int64_t ldc,
                          int64_t n, double alpha, double beta) {
  return guarded([&] {
    BCLX_REQUIRE(h && A && d_C, BCLX_ERR_INVALID, "null argument");
    BCLX_REQUIRE(A->dtype == BCLX_F64, BCLX_ERR_INVALID, "A must hold fp64 values");
    BCLX_REQUIRE(A->ptr_width == 4 || A->ptr_width == 8, BCLX_ERR_INVALID, "A->ptr_width must be 4 or 8");
    BCLX_REQUIRE(n >= 0 && ldb >= n && ldc >= n, BCLX_ERR_INVALID, "need ldb >= n and ldc >= n");
    BCLX_REQUIRE(A->nnz == 0 || (A->col_idx && A->vals && d_B), BCLX_ERR_INVALID, "null A arrays or B");
    set_device(h);
    scale_rows<double>(h, d_C, A->m, n, ldc, beta);
    if (A->ptr_width == 8) spmm_impl<int64_t, double, double>(h, A, d_B, ldb, d_C, ldc, n, alpha);
    else spmm_impl<int32_t, double, double>(h, A, d_B, ldb, d_C, ldc, n, alpha);
  });
}

}  // extern "C"
