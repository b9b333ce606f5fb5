// Training and inference kernels of the row-work predictor (C ABI: bclx_row_features, bclx_mlp_forward,
// bclx_mlp_grad, bclx_mlp_adam).
#include "internal.hpp"
#include "mlp.hpp"
#include "spgemm.hpp"

namespace bclx {

template <typename RP>
__global__ __launch_bounds__(256) void k_row_features(const RP* a_rp, const int* a_col, const RP* b_rp,
                                                      const int* b_col, int64_t m, int64_t n, int64_t kdim,
                                                      double avg_a, double avg_b, const int64_t* prod, float* feat) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  float f[kFeat];
  row_features<RP>(a_rp, a_col, b_rp, b_col, i, n, kdim, avg_a, avg_b, prod[i], f);
#pragma unroll
  for (int q = 0; q < kFeat; q++) feat[i * kFeat + q] = f[q];
}

// products per row (thread per row; used only by the feature export, not the hot path)
template <typename RP>
__global__ __launch_bounds__(256) void k_row_products(const RP* a_rp, const int* a_col, const RP* b_rp, int64_t m,
                                                      int64_t* prod) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  long long s = 0;
  for (RP p = a_rp[i]; p < a_rp[i + 1]; p++) {
    const int k = a_col[p];
    s += (long long)(b_rp[k + 1] - b_rp[k]);
  }
  prod[i] = s;
}

__global__ __launch_bounds__(256) void k_mlp_forward(int64_t nrows, const float* x, const float* params, float* y) {
  __shared__ float P[kNParam];
  for (int q = threadIdx.x; q < kNParam; q += blockDim.x) P[q] = params[q];
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  float xr[kFeat];
#pragma unroll
  for (int q = 0; q < kFeat; q++) xr[q] = x[r * kFeat + q];
  y[r] = mlp_forward_row(P, xr);
}

// Gradient of L = 0.5/N sum (y - t)^2. A block holds kGS samples: forward per sample (thread per sample), the
// per-sample activations and deltas go to LDS, then the block's weight-gradient partials are reductions over its
// samples (each thread owns a slice of the 5313 parameters), added to d_grads with one atomic per parameter.
constexpr int kGS = 64;

__global__ __launch_bounds__(256) void k_mlp_grad(int64_t nrows, const float* x, const float* t, const float* params,
                                                  float* grads, float* loss) {
  __shared__ float P[kNParam];
  // +1 padding: thread s walks row s, so rows must not start in the same LDS bank
  __shared__ float sx[kGS][kFeat + 1];
  __shared__ float sh1[kGS][kHid + 1];
  __shared__ float sh2[kGS][kHid + 1];
  __shared__ float sd1[kGS][kHid + 1];
  __shared__ float sd2[kGS][kHid + 1];
  __shared__ float sdy[kGS];
  for (int q = threadIdx.x; q < kNParam; q += blockDim.x) P[q] = params[q];
  __syncthreads();
  const float invn = 1.0f / (float)nrows;
  const int s = threadIdx.x;
  const int64_t r = (int64_t)blockIdx.x * kGS + s;
  if (s < kGS) {
    const bool ok = r < nrows;
    float xr[kFeat];
#pragma unroll
    for (int q = 0; q < kFeat; q++) {
      xr[q] = ok ? x[r * kFeat + q] : 0.f;
      sx[s][q] = xr[q];
    }
    float z1[kHid];
    for (int o = 0; o < kHid; o++) {
      float a = P[kOffB1 + o];
#pragma unroll
      for (int q = 0; q < kFeat; q++) a = fmaf(P[o * kFeat + q], xr[q], a);
      z1[o] = a;
      sh1[s][o] = a > 0.f ? a : 0.f;
    }
    float y = P[kOffB3];
    float z2[kHid];
    for (int o = 0; o < kHid; o++) {
      float a = P[kOffB2 + o];
      for (int q = 0; q < kHid; q++) a = fmaf(P[kOffW2 + o * kHid + q], sh1[s][q], a);
      z2[o] = a;
      const float h = a > 0.f ? a : 0.f;
      sh2[s][o] = h;
      y = fmaf(P[kOffW3 + o], h, y);
    }
    const float d = ok ? (y - t[r]) : 0.f;
    if (ok) atomicAdd(loss, 0.5f * d * d * invn);
    const float dy = d * invn;
    sdy[s] = dy;
    for (int o = 0; o < kHid; o++) sd2[s][o] = z2[o] > 0.f ? dy * P[kOffW3 + o] : 0.f;
    for (int q = 0; q < kHid; q++) {
      float a = 0.f;
      for (int o = 0; o < kHid; o++) a = fmaf(sd2[s][o], P[kOffW2 + o * kHid + q], a);
      sd1[s][q] = z1[q] > 0.f ? a : 0.f;
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < kNParam; p += blockDim.x) {
    float g = 0.f;
    if (p < kOffB1) {
      const int o = p / kFeat, q = p % kFeat;
      for (int u = 0; u < kGS; u++) g = fmaf(sd1[u][o], sx[u][q], g);
    } else if (p < kOffW2) {
      const int o = p - kOffB1;
      for (int u = 0; u < kGS; u++) g += sd1[u][o];
    } else if (p < kOffB2) {
      const int o = (p - kOffW2) / kHid, q = (p - kOffW2) % kHid;
      for (int u = 0; u < kGS; u++) g = fmaf(sd2[u][o], sh1[u][q], g);
    } else if (p < kOffW3) {
      const int o = p - kOffB2;
      for (int u = 0; u < kGS; u++) g += sd2[u][o];
    } else if (p < kOffB3) {
      const int o = p - kOffW3;
      for (int u = 0; u < kGS; u++) g = fmaf(sdy[u], sh2[u][o], g);
    } else {
      for (int u = 0; u < kGS; u++) g += sdy[u];
    }
    atomicAdd(&grads[p], g);
  }
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_zero_f32(float* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

template <typename RP>
static void features_impl(bclx_handle* h, const bclx_csr* A, const bclx_csr* B, float* feat, int64_t* prod) {
  const RP* arp = static_cast<const RP*>(A->row_ptr);
  const RP* brp = static_cast<const RP*>(B->row_ptr);
  const unsigned g = div_up(A->m, 256);
  if (A->m == 0) return;
  hipLaunchKernelGGL(k_row_products<RP>, dim3(g), dim3(256), 0, h->stream, arp, A->col_idx, brp, A->m, prod);
  BCLX_HIP(hipGetLastError());
  const double avg_a = A->m ? (double)A->nnz / (double)A->m : 0.0;
  const double avg_b = B->m ? (double)B->nnz / (double)B->m : 0.0;
  hipLaunchKernelGGL(k_row_features<RP>, dim3(g), dim3(256), 0, h->stream, arp, A->col_idx, brp, B->col_idx, A->m,
                     B->n, A->n, avg_a, avg_b, prod, feat);
  BCLX_HIP(hipGetLastError());
}

}  // namespace bclx

using namespace bclx;

extern "C" {

bclx_status bclx_row_features(bclx_handle* h, const bclx_csr* A, const bclx_csr* B, float* d_feat, int64_t* d_prod) {
  return guarded([&] {
    BCLX_REQUIRE(h && d_feat && d_prod, BCLX_ERR_INVALID, "null argument");
    check_pair(A, B);
    set_device(h);
    if (A->ptr_width == 8) features_impl<int64_t>(h, A, B, d_feat, d_prod);
    else features_impl<int32_t>(h, A, B, d_feat, d_prod);
  });
}

bclx_status bclx_mlp_forward(bclx_handle* h, int64_t nrows, const float* d_x, const float* d_params, float* d_y) {
  return guarded([&] {
    BCLX_REQUIRE(h && d_x && d_params && d_y && nrows >= 0, BCLX_ERR_INVALID, "bad argument");
    set_device(h);
    if (nrows == 0) return;
    hipLaunchKernelGGL(k_mlp_forward, dim3(div_up(nrows, 256)), dim3(256), 0, h->stream, nrows, d_x, d_params, d_y);
    BCLX_HIP(hipGetLastError());
  });
}

bclx_status bclx_mlp_grad(bclx_handle* h, int64_t nrows, const float* d_x, const float* d_t, const float* d_params,
                          float* d_grads, float* d_loss) {
  return guarded([&] {
    BCLX_REQUIRE(h && d_x && d_t && d_params && d_grads && d_loss && nrows > 0, BCLX_ERR_INVALID, "bad argument");
    set_device(h);
    hipLaunchKernelGGL(k_zero_f32, dim3(div_up(kNParam, 256)), dim3(256), 0, h->stream, d_grads, kNParam);
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, h->stream, d_loss, 1);
    hipLaunchKernelGGL(k_mlp_grad, dim3(div_up(nrows, kGS)), dim3(256), 0, h->stream, nrows, d_x, d_t, d_params,
                       d_grads, d_loss);
    BCLX_HIP(hipGetLastError());
  });
}

bclx_status bclx_mlp_adam(bclx_handle* h, float* d_params, const float* d_grads, float* d_m, float* d_v, int t,
                          float lr, float beta1, float beta2, float eps) {
  return guarded([&] {
    BCLX_REQUIRE(h && d_params && d_grads && d_m && d_v && t >= 1, BCLX_ERR_INVALID, "bad argument");
    set_device(h);
    const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(div_up(kNParam, 256)), dim3(256), 0, h->stream, d_params, d_grads, d_m, d_v,
                       kNParam, lr, beta1, beta2, eps, c1, c2);
    BCLX_HIP(hipGetLastError());
  });
}

}  // extern "C"
