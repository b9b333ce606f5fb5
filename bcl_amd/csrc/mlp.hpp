// Learned row-work predictor: per-row features of C = A * B and a 16 -> 64 -> 64 -> 1 ReLU MLP that predicts
// log2(1 + nnz(C_i)). Device functions shared by the stage pipeline (prediction inside the batched launch) and
// the C-ABI training / inference kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bclx {

constexpr int kFeat = 16;
constexpr int kHid = 64;
constexpr int kNParam = kHid * kFeat + kHid + kHid * kHid + kHid + kHid + 1;  // 5313
constexpr int kOffB1 = kHid * kFeat;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kFeatSample = 32;  // A entries sampled for the span features

__device__ inline float lg1p(double x) { return (float)log2(1.0 + (x > 0 ? x : 0.0)); }

// Features of row i (scaled to roughly [0, 1.5]). prod = products of the row (exact, from the analysis pass).
//  0 log2(1+nnz(A_i))/16          1 log2(1+prod)/32            2 log2(1+min(prod,n))/24
//  3 log2(1+max B row len)/24     4 log2(1+min B row len)/24    5 log2(1+n)/24
//  6 log2(1+union span)/24        7 log2(1+sum of spans)/32     8 log2(1+prod/union span)/16
//  9 log2(1+avg B row len)/16     10 log2(1+nnz(A_i)/avg A row)/8
//  11 log2(1+E_uniform)/24: expected distinct columns if prod draws fell uniformly in the union span
//  12 log2(1+E_uniform over n)/24 13 fraction of sampled B rows longer than 1
//  14 log2(1+sampled B length sum scaled to the whole row)/32   15 constant 1
template <typename RP>
__device__ inline void row_features(const RP* a_rp, const int* a_col, const RP* b_rp, const int* b_col, int64_t i,
                                    int64_t n, int64_t kdim, double avg_a, double avg_b, long long prod, float* f) {
  const RP a0 = a_rp[i], a1 = a_rp[i + 1];
  const long long na = (long long)(a1 - a0);
  long long mx = 0, mn = LLONG_MAX, span_sum = 0, lsum = 0, nlong = 0;
  int lo = INT_MAX, hi = -1;
  const long long step = na > kFeatSample ? na / kFeatSample : 1;
  int ns = 0;
  for (long long q = 0; q < na && ns < kFeatSample; q += step, ns++) {
    const int k = a_col[a0 + q];
    const RP b0 = b_rp[k], b1 = b_rp[k + 1];
    const long long len = (long long)(b1 - b0);
    lsum += len;
    if (len > mx) mx = len;
    if (len < mn) mn = len;
    if (len > 1) nlong++;
    if (len > 0) {
      const int c0 = b_col[b0], c1 = b_col[b1 - 1];
      if (c0 < lo) lo = c0;
      if (c1 > hi) hi = c1;
      span_sum += (long long)(c1 - c0 + 1);
    }
  }
  if (ns == 0) mn = 0;
  const double span = hi >= lo ? (double)(hi - lo + 1) : 1.0;
  const double dprod = (double)prod;
  const double e_span = span * (1.0 - exp(-dprod / span));
  const double e_n = (double)n * (1.0 - exp(-dprod / (double)(n > 0 ? n : 1)));
  const double lscaled = ns > 0 ? (double)lsum * (double)na / (double)ns : 0.0;
  f[0] = lg1p((double)na) / 16.f;
  f[1] = lg1p(dprod) / 32.f;
  f[2] = lg1p(dprod < (double)n ? dprod : (double)n) / 24.f;
  f[3] = lg1p((double)mx) / 24.f;
  f[4] = lg1p((double)mn) / 24.f;
  f[5] = lg1p((double)n) / 24.f;
  f[6] = lg1p(span) / 24.f;
  f[7] = lg1p((double)span_sum) / 32.f;
  f[8] = lg1p(dprod / span) / 16.f;
  f[9] = lg1p(avg_b) / 16.f;
  f[10] = lg1p((double)na / (avg_a > 0 ? avg_a : 1.0)) / 8.f;
  f[11] = lg1p(e_span) / 24.f;
  f[12] = lg1p(e_n) / 24.f;
  f[13] = ns > 0 ? (float)nlong / (float)ns : 0.f;
  f[14] = lg1p(lscaled) / 32.f;
  f[15] = 1.f;
  (void)kdim;
}

// Forward pass of one row with parameters in LDS (or any address space). Returns the raw output.
__device__ inline float mlp_forward_row(const float* __restrict__ P, const float* x) {
  float h1[kHid];
#pragma unroll 4
  for (int o = 0; o < kHid; o++) {
    float s = P[kOffB1 + o];
#pragma unroll
    for (int q = 0; q < kFeat; q++) s = fmaf(P[o * kFeat + q], x[q], s);
    h1[o] = s > 0.f ? s : 0.f;
  }
  float y = P[kOffB3];
  for (int o = 0; o < kHid; o++) {
    float s = P[kOffB2 + o];
#pragma unroll
    for (int q = 0; q < kHid; q++) s = fmaf(P[kOffW2 + o * kHid + q], h1[q], s);
    y = fmaf(P[kOffW3 + o], s > 0.f ? s : 0.f, y);
  }
  return y;
}

}  // namespace bclx
