// Deterministic block reductions to device scalars, shared by the BLAS-1,
// PCG and learned-smoother kernels.
//
// Each block reduces a fixed grid-stride slice (wavefront shuffles, then one
// LDS step across the block's waves) and stores a partial; the last block to
// finish (ticket counter) sums the partials in block order and writes the
// scalar. The block count depends only on N, so results are reproducible run
// to run. The counter is reset by the last block, so no memset is needed
// between calls (safe under stream capture).
#pragma once

#include "kcommon.hpp"

namespace libp_amd {

constexpr int RBS = 256;       // reduction block size
constexpr int RMAXB = 1024;    // max reduction blocks
constexpr int RNPART = 4;      // partial arrays available (fused multi-dot)

struct ReduceWS {
  double part[RNPART][RMAXB];
  unsigned int counter;
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

// Stores partials[k][blockIdx] for k < NV; returns true in thread 0 of the last block.
template <int NV>
__device__ inline bool finish_partials(ReduceWS* ws, const double (&v)[NV], double* s_w, bool& last) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double b = block_sum(v[k], s_w);
    if (threadIdx.x == 0) ws->part[k][blockIdx.x] = b;
  }
  __shared__ bool s_last;
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned ticket = atomicAdd(&ws->counter, 1u);
    s_last = (ticket == gridDim.x - 1);
  }
  __syncthreads();
  last = s_last;
  return last;
}

template <int NV>
__device__ inline void sum_partials(ReduceWS* ws, double (&out)[NV]) {
  // executed by the whole last block: each wave lane sums a strided subset in
  // a fixed order, then a fixed tree; deterministic for a given grid size.
  __shared__ double s_w[16];
  __threadfence();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double acc = 0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) acc += __builtin_nontemporal_load(&ws->part[k][b]);
    out[k] = block_sum(acc, s_w);
  }
  if (threadIdx.x == 0) ws->counter = 0;
}

// Partials without fence or ticket, for kernels that write large vectors: an
// agent-scope fence per block writes back the dirty L2, which in such kernels
// costs more than the reduction. A separate one-block finisher kernel then
// sums the partials (the kernel boundary makes them visible).
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
    double acc = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += ws->part[k][b];
    out[k] = block_sum(acc, s_w);
  }
}

}  // namespace libp_amd
