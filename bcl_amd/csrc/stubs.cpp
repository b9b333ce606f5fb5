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


}  // extern "C"
