// Error reporting for the C ABI.
#include <string>

#include "libp_amd.h"

namespace libp_amd {
static thread_local std::string g_last_error;
void set_last_error(const char* where, const char* what) {
  g_last_error = std::string(where) + ": " + what;
}
}  // namespace libp_amd

extern "C" const char* libp_last_error(void) { return libp_amd::g_last_error.c_str(); }

extern "C" const char* libp_status_string(int status) {
  switch (status) {
    case LIBP_OK: return "ok";
    case LIBP_ERR_ARG: return "invalid argument";
    case LIBP_ERR_HIP: return "HIP runtime error";
    case LIBP_ERR_SETUP: return "setup error";
    case LIBP_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}
