// Status strings and the per-thread last-error message of the C ABI.
#include <string>

#include "hsl2.h"

namespace hsl2 {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace hsl2

extern "C" int hsl2_version(void) { return 1; }

extern "C" const char* hsl2_last_error(void) { return hsl2::g_last_error.c_str(); }

extern "C" const char* hsl2_status_string(int status) {
  switch (status) {
    case HSL2_OK: return "ok";
    case HSL2_ERR_ARGUMENT: return "invalid argument";
    case HSL2_ERR_SHAPE: return "inconsistent shape";
    case HSL2_ERR_INDEX: return "index out of range";
    case HSL2_ERR_HIP: return "HIP runtime error";
    case HSL2_ERR_WORKSPACE: return "workspace too small";
    case HSL2_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}
