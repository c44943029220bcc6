// Shared helpers for the gfx950 kernels of libp_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "libp_amd.h"

namespace libp_amd {

// Records a message for libp_last_error() (defined in status.cpp).
void set_last_error(const char* where, const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Launch-status helper: every entry point returns through this.
inline int launch_status(const char* where) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_last_error(where, hipGetErrorString(e));
    return LIBP_ERR_HIP;
  }
  return LIBP_OK;
}

inline int arg_error(const char* where, const char* what) {
  set_last_error(where, what);
  return LIBP_ERR_ARG;
}

// 1-D grid for streaming kernels: enough blocks to cover N, capped so each
// thread walks a grid-stride loop (keeps launch count sane for N >> 10^8).
inline unsigned stream_blocks(long N, int bs, long cap = 65536) {
  long b = (N + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

// Wavefront (64-lane) sum of a double via DPP-friendly shuffles.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace libp_amd
