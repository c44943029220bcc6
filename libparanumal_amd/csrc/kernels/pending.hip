// Entry points whose kernels are not in this build yet: they report
// LIBP_ERR_UNSUPPORTED instead of computing anything.
#include "kcommon.hpp"

extern "C" int libp_mlp_linear_fwd(int, int, int, const float*, const unsigned short*, const float*, int, float*,
                                   void*) {
  libp_amd::set_last_error("mlp_linear_fwd", "not built");
  return LIBP_ERR_UNSUPPORTED;
}
extern "C" int libp_mlp_linear_bwd(int, int, int, const float*, const unsigned short*, const float*, const float*, int,
                                   float*, float*, float*, void*) {
  libp_amd::set_last_error("mlp_linear_bwd", "not built");
  return LIBP_ERR_UNSUPPORTED;
}
extern "C" int libp_adam_step(int, float*, unsigned short*, const float*, float*, float*, float, float, float, float,
                              int, void*) {
  libp_amd::set_last_error("adam_step", "not built");
  return LIBP_ERR_UNSUPPORTED;
}
