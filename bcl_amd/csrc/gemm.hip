// Dense GEMM C = beta * C + alpha * op(A) * op(B), row-major (bclx_gemm): the tile multiply BCL's dense gemm
// calls through cblas_gemm_wrapper_ (bcl/containers/algorithms/cblas_wrapper.hpp:16-40, called with
// alpha = beta = 1 from gemm_cowns, bcl/containers/algorithms/gemm.hpp:86-107).
//
// Three arithmetic modes on the matrix cores for fp32 operands (bclx_gemm), and an fp64 product (bclx_gemm_f64):
//   0  FP32: fp32 inputs, v_mfma_f32_32x32x2f32 (exact fp32 products, fp32 accumulation)
//   1  BF16: inputs rounded to bf16 (round to nearest even), v_mfma_f32_32x32x16_bf16, fp32 accumulation
//   2  split-BF16 3-term: x = x_hi + x_lo with x_hi = bf16(x), x_lo = bf16(x - x_hi); a*b is taken as
//      a_lo*b_hi + a_hi*b_lo + a_hi*b_hi (the a_lo*b_lo term, below 2^-16 relative, is dropped), fp32 accumulation
//
// Pipeline: a pack pass writes op(A) as Ap[M'][K'] and op(B) as Bp[N'][K'] (K contiguous, zero-padded to the block
// tile, converted / split for the bf16 modes), then the MFMA kernel runs on the packed operands with no bounds checks
// in its main loop. Block tile 128 x 128, four wavefronts of 64 x 64 (2 x 2 MFMA tiles of 32 x 32), the next K tile
// is loaded into registers while the current one is multiplied out of LDS. Blocks are ordered so that each XCD
// (blocks are dealt round-robin over the 8 XCDs) works on a compact patch of C tiles and shares their A and B panels
// in its L2.
#include <algorithm>
#include <type_traits>

#include "internal.hpp"

namespace bclx {

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E), so register arrays get constant indices.
template <int B, int E, typename F>
__device__ inline void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int kGThreads = 256;

// C tile of block b. XCD-aware: block b runs on XCD b % 8, so each XCD is given a contiguous run of the order below.
// The order walks groups of kTileGroup tile rows column by column, so the blocks an XCD runs at one time cover a
// near-square patch of C and share their A and B panels in that XCD's L2 (a row-major order would share one A
// panel and stream every B panel through L2; measured 8% slower on the BF16 8192^3 product). GROUPED = false keeps
// the row-major order.
constexpr int kTileGroup = 4;
template <bool GROUPED>
__device__ inline void gemm_tile_of(int b, int tiles_m, int tiles_n, int& tm, int& tn) {
  const int T = tiles_m * tiles_n;
  if (T % 8 == 0) b = (b % 8) * (T / 8) + (b / 8);
  if constexpr (GROUPED) {
    const int per_group = kTileGroup * tiles_n;
    const int first = (b / per_group) * kTileGroup;
    const int rows = tiles_m - first < kTileGroup ? tiles_m - first : kTileGroup;
    const int r = b % per_group;
    tm = first + r % rows;
    tn = r / rows;
  } else {
    tm = b / tiles_n;
    tn = b % tiles_n;
  }
}

__device__ inline uint16_t f32_to_bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays a (quiet) NaN
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf16_bits_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// dst[r * Kp + c] for r < Rp, c < Kp from src[r * rs + c * cs] (r < R, c < K), zero padding outside.
// MODE 0: fp32 copy; 1: bf16 (hi only); 2: bf16 hi and lo. 64 x 64 tiles staged through LDS: the reads run along
// the source's contiguous direction (64 consecutive floats per wavefront instruction) and each lane writes 8
// consecutive k of one row (16 bytes of bf16, 32 of fp32). Kp is a multiple of 32, so a chunk of 8 is wholly inside.
template <int MODE>
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ src, int64_t R, int64_t K, int64_t rs,
                                              int64_t cs, void* dst_hi, uint16_t* dst_lo, int64_t Rp, int64_t Kp) {
  __shared__ float tile[64][65];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  const bool rows_contig = cs != 1;  // source contiguous along r (a transposed operand)
  for (int q = w; q < 64; q += 4) {
    int64_t r, c;
    if (rows_contig) { r = r0 + lane; c = c0 + q; }
    else { r = r0 + q; c = c0 + lane; }
    const float v = (r < R && c < K) ? src[r * rs + c * cs] : 0.f;
    if (rows_contig) tile[lane][q] = v;
    else tile[q][lane] = v;
  }
  __syncthreads();
  for (int ch = tid; ch < 512; ch += 256) {
    const int rr = ch >> 3, cc = (ch & 7) * 8;
    const int64_t r = r0 + rr, c = c0 + cc;
    if (r >= Rp || c >= Kp) continue;
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = tile[rr][cc + i];
    if constexpr (MODE == 0) {
      u32x4* d = reinterpret_cast<u32x4*>(static_cast<float*>(dst_hi) + r * Kp + c);
      d[0] = u32x4{__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]), __float_as_uint(v[3])};
      d[1] = u32x4{__float_as_uint(v[4]), __float_as_uint(v[5]), __float_as_uint(v[6]), __float_as_uint(v[7])};
    } else {
      uint16_t hi[8], lo[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        hi[i] = f32_to_bf16_rne(v[i]);
        if constexpr (MODE == 2) lo[i] = f32_to_bf16_rne(v[i] - bf16_bits_to_f32(hi[i]));
      }
      auto pk = [](const uint16_t* x) {
        return u32x4{(uint32_t)x[0] | ((uint32_t)x[1] << 16), (uint32_t)x[2] | ((uint32_t)x[3] << 16),
                     (uint32_t)x[4] | ((uint32_t)x[5] << 16), (uint32_t)x[6] | ((uint32_t)x[7] << 16)};
      };
      *reinterpret_cast<u32x4*>(static_cast<uint16_t*>(dst_hi) + r * Kp + c) = pk(hi);
      if constexpr (MODE == 2) *reinterpret_cast<u32x4*>(dst_lo + r * Kp + c) = pk(lo);
    }
  }
}

template <int MODE>
struct GemmCfg;
template <>
struct GemmCfg<0> {
  using E = float;
  static constexpr int BK = 32, PAD = 4, KSTEP = 2;
};
template <>
struct GemmCfg<1> {
  using E = uint16_t;
  static constexpr int BK = 64, PAD = 8, KSTEP = 16;
};
template <>
struct GemmCfg<2> {
  using E = uint16_t;
  static constexpr int BK = 64, PAD = 8, KSTEP = 16;
};

// Block tile (64 * TM) x (64 * TN): four wavefronts in a 2 x 2 arrangement, each owning TM x TN MFMA tiles of
// 32 x 32. Larger wavefront tiles reuse each LDS fragment for more MFMAs (TM + TN fragment loads per k step feed
// TM * TN MFMAs).
template <int MODE, int TM, int TN>
__global__ __launch_bounds__(kGThreads, TM * TN > 4 ? 2 : 1) void k_gemm(const typename GemmCfg<MODE>::E* __restrict__ Ahi,
                                                    const typename GemmCfg<MODE>::E* __restrict__ Alo,
                                                    const typename GemmCfg<MODE>::E* __restrict__ Bhi,
                                                    const typename GemmCfg<MODE>::E* __restrict__ Blo, int64_t M,
                                                    int64_t N, int64_t Kp, float alpha, float beta, float* C,
                                                    int64_t ldc, int tiles_m, int tiles_n) {
  using E = typename GemmCfg<MODE>::E;
  constexpr int BK = GemmCfg<MODE>::BK, PAD = GemmCfg<MODE>::PAD, KSTEP = GemmCfg<MODE>::KSTEP;
  constexpr int BM = 64 * TM, BN = 64 * TN;
  constexpr int LDK = BK + PAD;
  constexpr int NSPLIT = MODE == 2 ? 2 : 1;
  constexpr int VPR = BK * (int)sizeof(E) / 16;   // 16-byte vectors per tile row
  constexpr int VPA = BM * VPR / kGThreads;       // vectors per thread, A tile
  constexpr int VPB = BN * VPR / kGThreads;       // vectors per thread, B tile
  __shared__ __align__(16) E As[NSPLIT][BM * LDK];
  __shared__ __align__(16) E Bs[NSPLIT][BN * LDK];

  int tm, tn;
  gemm_tile_of<true>(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  const E* __restrict__ Ag0 = Ahi + m0 * Kp;
  const E* __restrict__ Bg0 = Bhi + n0 * Kp;
  const E* __restrict__ Ag1 = MODE == 2 ? Alo + m0 * Kp : Ag0;
  const E* __restrict__ Bg1 = MODE == 2 ? Blo + n0 * Kp : Bg0;
  constexpr int EPV = 16 / (int)sizeof(E);  // elements per 16-byte vector
  // per-thread vector slots: row = (tid + i * 256) / VPR, column vector = (tid + i * 256) % VPR
  u32x4 ra0[VPA], rb0[VPB], ra1[VPA], rb1[VPB];
  auto gload = [&](int64_t k0) {
    static_for<0, VPA>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int v = tid + i * kGThreads;
      const int64_t o = (int64_t)(v / VPR) * Kp + k0 + (v % VPR) * EPV;
      ra0[i] = *reinterpret_cast<const u32x4*>(Ag0 + o);
      if constexpr (MODE == 2) ra1[i] = *reinterpret_cast<const u32x4*>(Ag1 + o);
    });
    static_for<0, VPB>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int v = tid + i * kGThreads;
      const int64_t o = (int64_t)(v / VPR) * Kp + k0 + (v % VPR) * EPV;
      rb0[i] = *reinterpret_cast<const u32x4*>(Bg0 + o);
      if constexpr (MODE == 2) rb1[i] = *reinterpret_cast<const u32x4*>(Bg1 + o);
    });
  };
  auto lstore = [&]() {
    static_for<0, VPA>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int v = tid + i * kGThreads;
      const int o = (v / VPR) * LDK + (v % VPR) * EPV;
      *reinterpret_cast<u32x4*>(&As[0][o]) = ra0[i];
      if constexpr (MODE == 2) *reinterpret_cast<u32x4*>(&As[NSPLIT - 1][o]) = ra1[i];
    });
    static_for<0, VPB>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int v = tid + i * kGThreads;
      const int o = (v / VPR) * LDK + (v % VPR) * EPV;
      *reinterpret_cast<u32x4*>(&Bs[0][o]) = rb0[i];
      if constexpr (MODE == 2) *reinterpret_cast<u32x4*>(&Bs[NSPLIT - 1][o]) = rb1[i];
    });
  };

  f32x16 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int c = 0; c < TN; c++)
#pragma unroll
      for (int v = 0; v < 16; v++) acc[a][c][v] = 0.f;

  const int nk = (int)(Kp / BK);
  gload(0);
  lstore();
  __syncthreads();
  const int fr = lane & 31, fk = lane >> 5;
  for (int kt = 0; kt < nk; kt++) {
    if (kt + 1 < nk) gload((int64_t)(kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += KSTEP) {
      if constexpr (MODE == 0) {
        float af[TM], bf[TN];
#pragma unroll
        for (int t = 0; t < TM; t++) af[t] = As[0][(wm * 32 * TM + t * 32 + fr) * LDK + ks + fk];
#pragma unroll
        for (int t = 0; t < TN; t++) bf[t] = Bs[0][(wn * 32 * TN + t * 32 + fr) * LDK + ks + fk];
#pragma unroll
        for (int a = 0; a < TM; a++)
#pragma unroll
          for (int c = 0; c < TN; c++) acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[a], bf[c], acc[a][c], 0, 0, 0);
      } else {
        bf16x8 ah[TM], bh[TN], al[TM], bl[TN];
#pragma unroll
        for (int t = 0; t < TM; t++) {
          ah[t] = *reinterpret_cast<const bf16x8*>(&As[0][(wm * 32 * TM + t * 32 + fr) * LDK + ks + fk * 8]);
          if constexpr (MODE == 2)
            al[t] = *reinterpret_cast<const bf16x8*>(&As[NSPLIT - 1][(wm * 32 * TM + t * 32 + fr) * LDK + ks + fk * 8]);
        }
#pragma unroll
        for (int t = 0; t < TN; t++) {
          bh[t] = *reinterpret_cast<const bf16x8*>(&Bs[0][(wn * 32 * TN + t * 32 + fr) * LDK + ks + fk * 8]);
          if constexpr (MODE == 2)
            bl[t] = *reinterpret_cast<const bf16x8*>(&Bs[NSPLIT - 1][(wn * 32 * TN + t * 32 + fr) * LDK + ks + fk * 8]);
        }
#pragma unroll
        for (int a = 0; a < TM; a++)
#pragma unroll
          for (int c = 0; c < TN; c++) {
            if constexpr (MODE == 2) {
              acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[c], acc[a][c], 0, 0, 0);
              acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[c], acc[a][c], 0, 0, 0);
            }
            acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[c], acc[a][c], 0, 0, 0);
          }
      }
    }
    __syncthreads();
    if (kt + 1 < nk) {
      lstore();
      __syncthreads();
    }
  }
  // accumulator layout of a 32 x 32 MFMA tile: lane l holds column l % 32, rows 8 * (v / 4) + 4 * (l / 32) + v % 4
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int c = 0; c < TN; c++) {
      const int64_t col = n0 + wn * 32 * TN + c * 32 + fr;
      if (col >= N) continue;
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int64_t row = m0 + wm * 32 * TM + a * 32 + 8 * (v / 4) + 4 * fk + (v % 4);
        if (row < M) {
          float* p = C + row * ldc + col;
          *p = beta == 0.f ? alpha * acc[a][c][v] : fmaf(beta, *p, alpha * acc[a][c][v]);
        }
      }
    }
}

// ---- FP64 GEMM on the matrix cores (cblas_gemm_wrapper_'s double overload -> cblas_dgemm,
// bcl/containers/algorithms/cblas_wrapper.hpp:29-40) ------------------------------------------------------------
// Packed operands as above (op(A) as [M'][K'], op(B) as [N'][K'], K contiguous, zero-padded), block tile 128 x 128
// of four wavefronts (2 x 2), each owning 4 x 4 tiles of 16 x 16 computed by v_mfma_f64_16x16x4f64 (fp64 products
// and sums). K tiles of 16 doubles are staged through LDS; the next one is loaded into registers meanwhile.

__global__ __launch_bounds__(256) void k_pack_f64(const double* __restrict__ src, int64_t R, int64_t K, int64_t rs,
                                                  int64_t cs, double* dst, int64_t Rp, int64_t Kp) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const bool rows_contig = cs != 1;
  for (int q = ty; q < 32; q += 8) {
    int64_t r, c;
    if (rows_contig) { r = r0 + tx; c = c0 + q; }
    else { r = r0 + q; c = c0 + tx; }
    const double v = (r < R && c < K) ? src[r * rs + c * cs] : 0.0;
    if (rows_contig) tile[tx][q] = v;
    else tile[q][tx] = v;
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int64_t r = r0 + q, c = c0 + tx;
    if (r < Rp && c < Kp) dst[r * Kp + c] = tile[q][tx];
  }
}

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
constexpr int kG64BK = 16;                // K per LDS tile
constexpr int kG64LDK = kG64BK + 2;       // padded row (bank spread)
constexpr int kG64T = 4;                  // 16 x 16 MFMA tiles per wavefront side
constexpr int kG64BM = 2 * 16 * kG64T;    // 128

__global__ __launch_bounds__(kGThreads) void k_gemm_f64(const double* __restrict__ Ap, const double* __restrict__ Bp,
                                                        int64_t M, int64_t N, int64_t Kp, double alpha, double beta,
                                                        double* C, int64_t ldc, int tiles_m, int tiles_n) {
  __shared__ __align__(16) double As[kG64BM * kG64LDK];
  __shared__ __align__(16) double Bs[kG64BM * kG64LDK];
  int tm, tn;  // row-major: the fp64 product is bound by its matrix-core rate, and the grouped order measured 4% slower
  gemm_tile_of<false>(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int64_t m0 = (int64_t)tm * kG64BM, n0 = (int64_t)tn * kG64BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  constexpr int VPR = kG64BK / 2;                 // 16-byte vectors per tile row
  constexpr int VPT = kG64BM * VPR / kGThreads;   // per thread, per operand
  const double* __restrict__ Ag = Ap + m0 * Kp;
  const double* __restrict__ Bg = Bp + n0 * Kp;
  f64x2 ra[VPT], rb[VPT];
  auto gload = [&](int64_t k0) {
    static_for<0, VPT>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int v = tid + i * kGThreads;
      const int64_t o = (int64_t)(v / VPR) * Kp + k0 + (v % VPR) * 2;
      ra[i] = *reinterpret_cast<const f64x2*>(Ag + o);
      rb[i] = *reinterpret_cast<const f64x2*>(Bg + o);
    });
  };
  auto lstore = [&]() {
    static_for<0, VPT>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int v = tid + i * kGThreads;
      const int o = (v / VPR) * kG64LDK + (v % VPR) * 2;
      *reinterpret_cast<f64x2*>(&As[o]) = ra[i];
      *reinterpret_cast<f64x2*>(&Bs[o]) = rb[i];
    });
  };
  f64x4 acc[kG64T][kG64T];
#pragma unroll
  for (int a = 0; a < kG64T; a++)
#pragma unroll
    for (int c = 0; c < kG64T; c++) acc[a][c] = f64x4{0.0, 0.0, 0.0, 0.0};
  const int nk = (int)(Kp / kG64BK);
  gload(0);
  lstore();
  __syncthreads();
  const int fr = lane & 15, fk = lane >> 4;  // operand lane layout: row (or column) fr, k offset fk
  for (int kt = 0; kt < nk; kt++) {
    if (kt + 1 < nk) gload((int64_t)(kt + 1) * kG64BK);
#pragma unroll
    for (int ks = 0; ks < kG64BK; ks += 4) {
      double af[kG64T], bf[kG64T];
#pragma unroll
      for (int t = 0; t < kG64T; t++) {
        af[t] = As[(wm * 16 * kG64T + t * 16 + fr) * kG64LDK + ks + fk];
        bf[t] = Bs[(wn * 16 * kG64T + t * 16 + fr) * kG64LDK + ks + fk];
      }
#pragma unroll
      for (int a = 0; a < kG64T; a++)
#pragma unroll
        for (int c = 0; c < kG64T; c++) acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[c], acc[a][c], 0, 0, 0);
    }
    __syncthreads();
    if (kt + 1 < nk) {
      lstore();
      __syncthreads();
    }
  }
  // accumulator layout of a 16 x 16 fp64 tile: lane l holds column l % 16, rows l / 16 + 4 * v
  static_for<0, kG64T>([&](auto ac) {
    constexpr int a = decltype(ac)::value;
    static_for<0, kG64T>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      const int64_t col = n0 + wn * 16 * kG64T + c * 16 + fr;
      static_for<0, 4>([&](auto vc) {
        constexpr int v = decltype(vc)::value;
        const int64_t row = m0 + wm * 16 * kG64T + a * 16 + fk + 4 * v;
        if (row < M && col < N) {
          double* p = C + row * ldc + col;
          *p = beta == 0.0 ? alpha * acc[a][c][v] : fma(beta, *p, alpha * acc[a][c][v]);
        }
      });
    });
  });
}

static void gemm_f64_run(bclx_handle* h, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                         const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                         int64_t ldc) {
  const int64_t Mp = (m + kG64BM - 1) / kG64BM * kG64BM, Np = (n + kG64BM - 1) / kG64BM * kG64BM;
  const int64_t Kp = std::max<int64_t>(32, (k + 31) / 32 * 32);  // whole pack tiles and whole K tiles
  const size_t a_bytes = (size_t)Mp * Kp * 8, off_b = (a_bytes + 255) / 256 * 256;
  h->scratch.reserve(off_b + (size_t)Np * Kp * 8 + 256);
  double* Ap = h->scratch.as<double>();
  double* Bp = reinterpret_cast<double*>(h->scratch.as<char>() + off_b);
  hipLaunchKernelGGL(k_pack_f64, dim3((unsigned)(Kp / 32), (unsigned)(Mp / 32)), dim3(256), 0, h->stream, A, m, k,
                     ta ? (int64_t)1 : lda, ta ? lda : (int64_t)1, Ap, Mp, Kp);
  BCLX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pack_f64, dim3((unsigned)(Kp / 32), (unsigned)(Np / 32)), dim3(256), 0, h->stream, B, n, k,
                     tb ? ldb : (int64_t)1, tb ? (int64_t)1 : ldb, Bp, Np, Kp);
  BCLX_HIP(hipGetLastError());
  const int tiles_m = (int)(Mp / kG64BM), tiles_n = (int)(Np / kG64BM);
  hipLaunchKernelGGL(k_gemm_f64, dim3((unsigned)(tiles_m * tiles_n)), dim3(kGThreads), 0, h->stream, Ap, Bp, m, n, Kp,
                     alpha, beta, C, ldc, tiles_m, tiles_n);
  BCLX_HIP(hipGetLastError());
}

// FP64 reference: C += A * B (fp32 inputs widened exactly), 16 x 16 tiles, row-major. Error oracle only.
__global__ __launch_bounds__(256) void k_gemm_f64_ref(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                                                      const float* B, int64_t ldb, double* C, int64_t ldc) {
  __shared__ double as[16][17], bs[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t row = (int64_t)blockIdx.y * 16 + ty, col = (int64_t)blockIdx.x * 16 + tx;
  double s = 0.0;
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    as[ty][tx] = (row < M && k0 + tx < K) ? (double)A[row * lda + k0 + tx] : 0.0;
    bs[ty][tx] = (k0 + ty < K && col < N) ? (double)B[(k0 + ty) * ldb + col] : 0.0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++) s = fma(as[ty][q], bs[q][tx], s);
    __syncthreads();
  }
  if (row < M && col < N) C[row * ldc + col] += s;
}

// Wavefront tile (TM x TN MFMA tiles) per mode: BF16 uses 4 x 2 (block 256 x 128) when M is large enough to fill the
// GPU with such blocks; FP32 and split-BF16 (twice the LDS) keep 2 x 2 (block 128 x 128).
template <int MODE, int TM, int TN>
static void gemm_run(bclx_handle* h, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha, const float* A,
                     int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc) {
  using E = typename GemmCfg<MODE>::E;
  constexpr int BK = GemmCfg<MODE>::BK, BM = 64 * TM, BN = 64 * TN;
  const int64_t Mp = (m + BM - 1) / BM * BM, Np = (n + BN - 1) / BN * BN;
  const int64_t Kp = std::max<int64_t>(BK, (k + BK - 1) / BK * BK);
  const size_t a_bytes = (size_t)Mp * Kp * sizeof(E), b_bytes = (size_t)Np * Kp * sizeof(E);
  const size_t lo_a = MODE == 2 ? (size_t)Mp * Kp * 2 : 0, lo_b = MODE == 2 ? (size_t)Np * Kp * 2 : 0;
  const size_t off_b = (a_bytes + 255) / 256 * 256, off_la = off_b + (b_bytes + 255) / 256 * 256;
  const size_t off_lb = off_la + (lo_a + 255) / 256 * 256, total = off_lb + lo_b + 256;
  h->scratch.reserve(total);
  char* ws = h->scratch.as<char>();
  E* Ap = reinterpret_cast<E*>(ws);
  E* Bp = reinterpret_cast<E*>(ws + off_b);
  uint16_t* Al = MODE == 2 ? reinterpret_cast<uint16_t*>(ws + off_la) : nullptr;
  uint16_t* Bl = MODE == 2 ? reinterpret_cast<uint16_t*>(ws + off_lb) : nullptr;
  // op(A)[i][kk] = ta ? A[kk * lda + i] : A[i * lda + kk]; Bp[j][kk] = op(B)[kk][j] = tb ? B[j * ldb + kk] : B[kk * ldb + j]
  dim3 blk(256);
  hipLaunchKernelGGL(k_pack<MODE>, dim3((unsigned)div_up(Kp, 64), (unsigned)div_up(Mp, 64)), blk, 0, h->stream, A, m, k,
                     ta ? (int64_t)1 : lda, ta ? lda : (int64_t)1, (void*)Ap, Al, Mp, Kp);
  BCLX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pack<MODE>, dim3((unsigned)div_up(Kp, 64), (unsigned)div_up(Np, 64)), blk, 0, h->stream, B, n, k,
                     tb ? ldb : (int64_t)1, tb ? (int64_t)1 : ldb, (void*)Bp, Bl, Np, Kp);
  BCLX_HIP(hipGetLastError());
  const int tiles_m = (int)(Mp / BM), tiles_n = (int)(Np / BN);
  hipLaunchKernelGGL((k_gemm<MODE, TM, TN>), dim3((unsigned)(tiles_m * tiles_n)), dim3(kGThreads), 0, h->stream, Ap,
                     (const E*)Al, Bp, (const E*)Bl, m, n, Kp, alpha, beta, C, ldc, tiles_m, tiles_n);
  BCLX_HIP(hipGetLastError());
}

template <int MODE>
static void gemm_impl(bclx_handle* h, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha, const float* A,
                      int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc) {
  const bool big = MODE == 1 && ((m + 255) / 256) * ((n + 127) / 128) >= 2 * (int64_t)h->num_cus;
  if (big) gemm_run<MODE, 4, 2>(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  else gemm_run<MODE, 2, 2>(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

}  // namespace bclx

using namespace bclx;

extern "C" {

bclx_status bclx_gemm(bclx_handle* h, int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k, float alpha,
                      const float* d_A, int64_t lda, const float* d_B, int64_t ldb, float beta, float* d_C,
                      int64_t ldc, int32_t mode) {
  return guarded([&] {
    BCLX_REQUIRE(h && d_C, BCLX_ERR_INVALID, "null handle or C");
    BCLX_REQUIRE(m >= 0 && n >= 0 && k >= 0, BCLX_ERR_INVALID, "negative size");
    BCLX_REQUIRE(mode >= 0 && mode <= 2, BCLX_ERR_INVALID, "mode must be 0 (fp32), 1 (bf16) or 2 (split-bf16)");
    BCLX_REQUIRE(ldc >= n, BCLX_ERR_INVALID, "ldc < n");
    BCLX_REQUIRE(lda >= (trans_a ? m : k) && ldb >= (trans_b ? k : n), BCLX_ERR_INVALID, "lda / ldb too small");
    BCLX_REQUIRE(k == 0 || (d_A && d_B), BCLX_ERR_INVALID, "null A or B");
    set_device(h);
    if (m == 0 || n == 0) return;  // k == 0 packs zeros, so C = beta * C
    const bool ta = trans_a != 0, tb = trans_b != 0;
    switch (mode) {
      case 0: gemm_impl<0>(h, ta, tb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc); break;
      case 1: gemm_impl<1>(h, ta, tb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc); break;
      default: gemm_impl<2>(h, ta, tb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc); break;
    }
  });
}

bclx_status bclx_gemm_f64(bclx_handle* h, int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k,
                          double alpha, const double* d_A, int64_t lda, const double* d_B, int64_t ldb, double beta,
                          double* d_C, int64_t ldc) {
  return guarded([&] {
    BCLX_REQUIRE(h && d_C, BCLX_ERR_INVALID, "null handle or C");
    BCLX_REQUIRE(m >= 0 && n >= 0 && k >= 0, BCLX_ERR_INVALID, "negative size");
    BCLX_REQUIRE(ldc >= n, BCLX_ERR_INVALID, "ldc < n");
    BCLX_REQUIRE(lda >= (trans_a ? m : k) && ldb >= (trans_b ? k : n), BCLX_ERR_INVALID, "lda / ldb too small");
    BCLX_REQUIRE(k == 0 || (d_A && d_B), BCLX_ERR_INVALID, "null A or B");
    set_device(h);
    if (m == 0 || n == 0) return;
    gemm_f64_run(h, trans_a != 0, trans_b != 0, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
  });
}

bclx_status bclx_gemm_f64_ref(bclx_handle* h, int64_t m, int64_t n, int64_t k, const float* d_A, int64_t lda,
                              const float* d_B, int64_t ldb, double* d_C, int64_t ldc) {
  return guarded([&] {
    BCLX_REQUIRE(h && d_A && d_B && d_C && m >= 0 && n >= 0 && k >= 0, BCLX_ERR_INVALID, "bad argument");
    set_device(h);
    if (m == 0 || n == 0) return;
    hipLaunchKernelGGL(k_gemm_f64_ref, dim3(div_up(n, 16), div_up(m, 16)), dim3(256), 0, h->stream, m, n, k, d_A,
                       lda, d_B, ldb, d_C, ldc);
    BCLX_HIP(hipGetLastError());
  });
}

}  // extern "C"
