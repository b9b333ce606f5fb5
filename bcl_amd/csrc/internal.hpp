// Internal runtime of the bclx library: handle, error reporting, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bclx.h"

namespace bclx {

void set_error(const std::string& msg);

struct Error {
  bclx_status status;
  std::string msg;
};

#define BCLX_HIP(call)                                                                                    \
  do {                                                                                                     \
    hipError_t e_ = (call);                                                                                \
    if (e_ != hipSuccess) {                                                                                \
      throw ::bclx::Error{e_ == hipErrorOutOfMemory ? BCLX_ERR_ALLOC : BCLX_ERR_HIP,                      \
                          std::string(#call) + ": " + hipGetErrorString(e_) + " at " + __FILE__ + ":" +   \
                              std::to_string(__LINE__)};                                                   \
    }                                                                                                      \
  } while (0)

#define BCLX_REQUIRE(cond, status, msg)                       \
  do {                                                        \
    if (!(cond)) throw ::bclx::Error{(status), (msg)};        \
  } while (0)

// The device allocator of the calling thread's current handle (bclx_set_allocator; hipMalloc / hipFree when unset).
// Every entry point selects its handle's allocator through set_device() before it allocates.
struct AllocFns {
  bclx_alloc_fn alloc = nullptr;
  bclx_free_fn free = nullptr;
  void* user = nullptr;
};
extern thread_local AllocFns g_alloc;

// Device memory from the current allocator; freed with the allocator it came from.
void* dev_alloc(size_t bytes, const AllocFns& how);
void dev_free(void* p, const AllocFns& how);

// Device buffer owned by the library. Move-only; remembers the allocator that made it.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  AllocFns how;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), how(o.how) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; how = o.how; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n) {
    release();
    if (n == 0) n = 256;
    how = g_alloc;
    p = dev_alloc(n, how);
    bytes = n;
  }
  // grow-only
  void reserve(size_t n) { if (n > bytes) alloc(n); }
  void release() {
    if (p) dev_free(p, how);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Bump allocator over one device buffer: sub-buffers are 256-byte aligned. The buffer is the arena's own (allocated
// at commit), or caller memory given to use_external (nothing is allocated; commit fails when it is too small).
struct Arena {
  DevBuf buf;
  char* ext = nullptr;
  size_t ext_bytes = 0;
  size_t used = 0;
  size_t reserve(size_t bytes) {
    size_t off = (used + 255) & ~size_t(255);
    used = off + bytes;
    return off;
  }
  void use_external(void* p, size_t bytes) {
    ext = static_cast<char*>(p);
    ext_bytes = bytes;
  }
  void commit() {
    if (ext) {
      BCLX_REQUIRE(used <= ext_bytes, BCLX_ERR_OVERFLOW,
                   "workspace of " + std::to_string(used) + " bytes exceeds the " + std::to_string(ext_bytes) +
                       "-byte scratch");
    } else {
      buf.reserve(used);
    }
  }
  template <typename T> T* at(size_t off) const {
    return reinterpret_cast<T*>((ext ? ext : static_cast<char*>(buf.p)) + off);
  }
};

}  // namespace bclx

struct bclx_handle {
  int device = 0;
  bclx::AllocFns alloc;  // device allocator of this handle (bclx_set_allocator)
  hipStream_t stream = nullptr;
  int num_cus = 256;
  size_t lds_per_block = 65536;
  // scratch reused across calls (grow-only)
  bclx::DevBuf scratch;
  bclx::DevBuf dense_slots;  // heavy-row dense accumulators, kept zeroed between uses
  size_t dense_slot_doubles = 0;
  int dense_nslots = 0;
  bclx::DevBuf pinned_dummy;
  int64_t* h_pinned = nullptr;  // small pinned host buffer for counts
};

namespace bclx {

// Wraps a C-ABI body: converts exceptions to status codes and records the message.
template <typename F>
bclx_status guarded(F&& f) {
  try {
    f();
    return BCLX_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.status;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return BCLX_ERR_ALLOC;
  } catch (const std::exception& e) {
    set_error(e.what());
    return BCLX_ERR_INVALID;
  }
}

inline void set_device(bclx_handle* h) {
  BCLX_HIP(hipSetDevice(h->device));
  g_alloc = h->alloc;
}

inline unsigned div_up(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Looks up, once per device and with that device current, the address of the fixed scratch array that
// bclx_spgemm_noalloc works in.
void init_noalloc_scratch(int device);

// Zeroed heavy-row accumulators: nslots x ndoubles, grown on demand, zeroed once when (re)allocated.
void ensure_dense_slots(bclx_handle* h, int nslots, size_t ndoubles);

}  // namespace bclx
