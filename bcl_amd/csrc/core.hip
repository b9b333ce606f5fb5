// Handle lifecycle, error reporting and version of the bclx C ABI.
#include <memory>

#include "internal.hpp"

namespace bclx {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

thread_local AllocFns g_alloc;

void* dev_alloc(size_t bytes, const AllocFns& how) {
  void* p = nullptr;
  if (how.alloc) {
    p = how.alloc(bytes, how.user);
    BCLX_REQUIRE(p != nullptr, BCLX_ERR_ALLOC,
                 "the handle's allocator returned NULL for " + std::to_string(bytes) + " bytes");
  } else {
    BCLX_HIP(hipMalloc(&p, bytes));
  }
  return p;
}

void dev_free(void* p, const AllocFns& how) {
  if (how.free) how.free(p, how.user);
  else (void)hipFree(p);
}
}  // namespace bclx

using namespace bclx;

extern "C" {

const char* bclx_last_error(void) { return g_last_error.c_str(); }

const char* bclx_version(void) { return "bclx 0.1 (gfx950)"; }

bclx_status bclx_create(int device, void* stream, bclx_handle** out) {
  return guarded([&] {
    BCLX_REQUIRE(out != nullptr, BCLX_ERR_INVALID, "out is null");
    int ndev = 0;
    BCLX_HIP(hipGetDeviceCount(&ndev));
    BCLX_REQUIRE(device >= 0 && device < ndev, BCLX_ERR_INVALID,
                 "device " + std::to_string(device) + " not present (" + std::to_string(ndev) + " devices)");
    BCLX_HIP(hipSetDevice(device));
    std::unique_ptr<bclx_handle> h(new bclx_handle());
    h->device = device;
    h->stream = static_cast<hipStream_t>(stream);
    int cus = 0, lds = 0;
    BCLX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    BCLX_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    h->num_cus = cus > 0 ? cus : 256;
    h->lds_per_block = lds > 0 ? (size_t)lds : 65536;
    init_noalloc_scratch(device);
    *out = h.release();
  });
}

bclx_status bclx_destroy(bclx_handle* h) {
  return guarded([&] {
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    delete h;
  });
}

bclx_status bclx_set_allocator(bclx_handle* h, bclx_alloc_fn alloc, bclx_free_fn free_fn, void* user) {
  return guarded([&] {
    BCLX_REQUIRE(h != nullptr, BCLX_ERR_INVALID, "null handle");
    BCLX_REQUIRE((alloc == nullptr) == (free_fn == nullptr), BCLX_ERR_INVALID,
                 "give both allocator functions, or neither (hipMalloc / hipFree)");
    h->alloc = AllocFns{alloc, free_fn, user};
  });
}

bclx_status bclx_set_stream(bclx_handle* h, void* stream) {
  return guarded([&] {
    BCLX_REQUIRE(h != nullptr, BCLX_ERR_INVALID, "null handle");
    h->stream = static_cast<hipStream_t>(stream);
  });
}

bclx_status bclx_synchronize(bclx_handle* h) {
  return guarded([&] {
    BCLX_REQUIRE(h != nullptr, BCLX_ERR_INVALID, "null handle");
    set_device(h);
    BCLX_HIP(hipStreamSynchronize(h->stream));
  });
}

}  // extern "C"
