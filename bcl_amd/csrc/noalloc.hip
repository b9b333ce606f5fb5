// C = A * B in one call that allocates no device memory (bclx_spgemm_noalloc), for callers that cannot allocate:
// code under HIP graph capture, serving loops that keep memory fixed, harnesses that forbid hipMalloc.
//
// It runs the kernels of bclx_spgemm (symbolic pass, then numeric pass into the caller's C) with their per-row
// workspace (products, counts, offsets, bin lists, scan partials) laid out in `scratch`, a fixed device array of
// kScratchBytes that each GPU holds from the library's load until the process ends. Its address is looked up once
// per device, with that device current, when a handle is made on it (bclx_create); host code only ever passes that
// looked-up address, never the array's own name. Every call on one GPU shares the array and leaves its contents
// behind: each pass clears the counters it uses before running, and calls that overlap on one GPU (two host threads,
// or two streams) race on it.
#include <mutex>

#include "internal.hpp"
#include "spgemm.hpp"

namespace bclx {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB: about 600 K rows of A per call
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Called by bclx_create with `device` current.
void init_noalloc_scratch(int device) {
  BCLX_REQUIRE(device >= 0 && device < kMaxDevices, BCLX_ERR_UNSUPPORTED,
               "devices numbered above " + std::to_string(kMaxDevices - 1) + " are not supported");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  BCLX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  BCLX_REQUIRE(p != nullptr, BCLX_ERR_HIP, "no device address for the scratch array");
  g_scratch_ptr[device] = p;
}

static void* scratch_on(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  void* p = g_scratch_ptr[device];
  BCLX_REQUIRE(p != nullptr, BCLX_ERR_INVALID, "the scratch of this device was not set up (make the handle with bclx_create)");
  return p;
}

template <typename T, typename RP>
static void spgemm_noalloc(bclx_handle* h, const bclx_csr* A, const bclx_csr* B, bclx_csr* C) {
  const int64_t cap = C->nnz;
  TwoPhase<T, RP> W(h, make_tile<T, RP>(A, B), BinTuning{}, scratch_on(h->device), kScratchBytes);
  const int64_t nnz = W.symbolic();
  if (nnz > cap) {
    C->nnz = nnz;  // what a retry needs
    throw Error{BCLX_ERR_OVERFLOW, "nnz(C) = " + std::to_string(nnz) + " exceeds the " + std::to_string(cap) +
                                       " entries C holds (C->nnz now says how many are needed)"};
  }
  W.numeric(C->col_idx, static_cast<T*>(C->vals), cap, C->row_ptr, C->ptr_width);
  BCLX_HIP(hipStreamSynchronize(h->stream));
  C->nnz = nnz;
}

}  // namespace bclx

using namespace bclx;

extern "C" {

bclx_status bclx_spgemm_noalloc(bclx_handle* h, const bclx_csr* A, const bclx_csr* B, bclx_csr* C) {
  return guarded([&] {
    BCLX_REQUIRE(h && C, BCLX_ERR_INVALID, "null handle or C");
    check_pair(A, B);
    BCLX_REQUIRE(C->row_ptr != nullptr, BCLX_ERR_INVALID, "C->row_ptr is null");
    BCLX_REQUIRE(C->ptr_width == 4 || C->ptr_width == 8, BCLX_ERR_INVALID, "C->ptr_width must be 4 or 8");
    BCLX_REQUIRE(C->nnz >= 0, BCLX_ERR_INVALID, "C->nnz (the capacity of C's arrays) is negative");
    BCLX_REQUIRE(C->nnz == 0 || (C->col_idx && C->vals), BCLX_ERR_INVALID, "C->col_idx / C->vals are null");
    BCLX_REQUIRE(C->dtype == A->dtype, BCLX_ERR_INVALID, "C dtype must match A and B");
    BCLX_REQUIRE(B->n <= (int64_t)kHRWords * 32, BCLX_ERR_UNSUPPORTED,
                 "B has more than 2^19 columns: its heavy rows need the accumulators bclx_spgemm allocates");
    const int64_t rows = (A->m + kRowBlock - 1) / kRowBlock * kRowBlock;
    BCLX_REQUIRE(rows * 112 + (int64_t)(1 << 16) <= (int64_t)kScratchBytes, BCLX_ERR_OVERFLOW,
                 "A has " + std::to_string(A->m) + " rows: more than the " + std::to_string(kScratchBytes >> 20) +
                     " MiB scratch of bclx_spgemm_noalloc holds (about " +
                     std::to_string(((int64_t)kScratchBytes - (1 << 16)) / 112) + " rows)");
    set_device(h);
    C->m = A->m;
    C->n = B->n;
    const bool f64 = A->dtype == BCLX_F64, w8 = A->ptr_width == 8;
    if (!f64 && !w8) spgemm_noalloc<float, int32_t>(h, A, B, C);
    else if (!f64) spgemm_noalloc<float, int64_t>(h, A, B, C);
    else if (!w8) spgemm_noalloc<double, int32_t>(h, A, B, C);
    else spgemm_noalloc<double, int64_t>(h, A, B, C);
  });
}

}  // extern "C"
