// Entry points declared in bclx.h whose implementation is not in this build yet: they report
// BCLX_ERR_UNSUPPORTED with a message instead of computing anything.
#include "../../include/bclx.h"

#include <string>

namespace bclx {
void set_error(const std::string& msg);
}

#define BCLX_STUB(name)                                           \
  bclx::set_error(std::string(name) + ": not implemented yet"); \
  return BCLX_ERR_UNSUPPORTED

extern "C" {

bclx_status bclx_accumulate_tiles(bclx_handle*, int, const bclx_csr*, bclx_csr*) { BCLX_STUB("bclx_accumulate_tiles"); }
bclx_status bclx_spmm(bclx_handle*, const bclx_csr*, const void*, int32_t, int64_t, float*, int64_t, int64_t, float,
                      float) { BCLX_STUB("bclx_spmm"); }
bclx_status bclx_gemm(bclx_handle*, int32_t, int32_t, int64_t, int64_t, int64_t, float, const float*, int64_t,
                      const float*, int64_t, float, float*, int64_t, int32_t) { BCLX_STUB("bclx_gemm"); }
bclx_status bclx_gemm_f64_ref(bclx_handle*, int64_t, int64_t, int64_t, const float*, int64_t, const float*, int64_t,
                              double*, int64_t) { BCLX_STUB("bclx_gemm_f64_ref"); }

}  // extern "C"
