// Deterministic block reductions to device scalars, shared by the BLAS-1,
// PCG and learned-smoother kernels.
//
// A reducing kernel has each block reduce a fixed grid-stride slice
// (wavefront shuffles, then one LDS step across the block's waves) and store
// one partial per block, with no fence and no ticket. A one-block finisher
// kernel launched next sums the partials in block order and writes the
// scalars; the kernel boundary makes the partials visible. The block count
// depends only on N, so results are reproducible run to run. An agent-scope
// fence per block followed by a last-block ticket, the usual one-kernel form,
// cost 25-40% of a streaming reduction's time on this GPU (it writes back and
// invalidates the L2 of each XCD once per block). Nothing needs clearing
// between calls, so the kernels are safe under stream capture.
#pragma once

#include "kcommon.hpp"

namespace libp_amd {

constexpr int RBS = 256;       // reduction block size
constexpr int RMAXB = 1024;    // max reduction blocks
constexpr int RNPART = 4;      // partial arrays available (fused multi-dot)

struct ReduceWS {
  double part[RNPART][RMAXB];
};

inline unsigned reduce_blocks(long N) {
  long b = (N + RBS * 4 - 1) / (RBS * 4);
  if (b < 1) b = 1;
  if (b > RMAXB) b = RMAXB;
  return (unsigned)b;
}

// Block-level sum; result valid in thread 0.
__device__ inline double block_sum(double v, double* s_w) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s_w[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r += s_w[i];
  }
  __syncthreads();
  return r;
}

// Sum of x[i] for i = threadIdx.x, threadIdx.x + blockDim.x, ... below n,
// added from zero in that order: the order of the plain strided loop, so the
// result is the same bit for bit, but with eight loads in flight (at clamped
// indices; entries past n add nothing) instead of a chain of dependent loads.
// A one-block finisher over 73,429 partials (48^3 PCG gather-update) took
// 34 us with the plain loop. U loads per thread per round (more for the
// finishers whose partials come from other XCDs' blocks: each round is a far
// load's latency).
template <int U = 8>
__device__ inline double strided_sum(const double* __restrict__ x, const int n) {
  double t = 0;
  const int B = blockDim.x;
  for (int i0 = threadIdx.x; i0 < n; i0 += U * B) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = x[min(i0 + u * B, n - 1)];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * B < n) t += v[u];
  }
  return t;
}

// One partial per block for each of NV sums: ws->part[k][blockIdx.x].
template <int NV>
__device__ inline void store_block_partials(ReduceWS* ws, const double (&v)[NV], double* s_w) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double b = block_sum(v[k], s_w);
    if (threadIdx.x == 0) ws->part[k][blockIdx.x] = b;
  }
}

// Finisher side (one block): the nb partials of each slot summed in a fixed
// order; valid in thread 0.
template <int NV>
__device__ inline void sum_block_partials(const ReduceWS* ws, const int nb, double (&out)[NV], double* s_w) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    out[k] = block_sum(strided_sum(ws->part[k], nb), s_w);
  }
}

}  // namespace libp_amd
