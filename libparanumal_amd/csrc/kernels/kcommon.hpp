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

// Position of copy k of a CSR row [a, a + len), clamped into the row: the
// gather kernels load the first eight copies of a row unconditionally (copies
// past the row's end re-read its last copy, which is not summed), so all the
// loads are independent and in flight together. A load under `k < len` is
// compiled as a branch that waits for each load before the next one is issued.
// A row with len == 0 (and a thread past the last row, a = len = 0) reads an
// earlier valid entry; the arrays hold at least one entry.
__device__ inline int row_copy(const int a, const int len, const int k) { return min(a + k, max(a + len - 1, 0)); }

// The largest v over the 64 lanes of a wavefront, as a scalar: a loop over
// k < wave_max_uniform(len) is a uniform branch, so the loads it guards need
// not wait for one another. Every lane of the wavefront must be active.
__device__ inline int wave_max_uniform(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

// Wavefront (64-lane) sum of a double via DPP-friendly shuffles.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace libp_amd
