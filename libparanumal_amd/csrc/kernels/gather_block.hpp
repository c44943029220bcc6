// Row-block staging for the gather kernels (ogs Trans gather and the kernels
// fused with it).
//
// The reference stages the copies of a block of rows through shared memory
// (libs/ogs/okl/ogsKernels.okl:86-122). Here a workgroup of GB_THREADS
// threads takes GB_ROWS consecutive rows: their copies are one contiguous
// range of the CSR, so the workgroup loads every copy exactly once (the
// column-index loads coalesced, one value load per copy, all in flight
// before any is used), stages the values in LDS, and each thread then sums
// its two rows from LDS from zero in ascending order, the arithmetic of a
// thread-per-row gather, so results are bit-identical to it. A thread per
// row instead issues one load slot per copy position of the longest row of
// its wavefront; on the 48^3, N=7 box this staging took the FP64 gather from
// 340 to ~200 us. A block with more than GB_CAP copies (the N=1 level of a
// hex mesh averages close to eight copies per row, ~4000 per block) is staged
// in chunks of GB_CAP copies, in ascending order; each row adds the copies of
// each chunk it overlaps in turn, so its sum keeps the same order. GB_CAP
// sets the LDS per workgroup: 2048 copies (16 KB in FP64) keeps eight
// workgroups per CU resident, where a 4096 cap took the 48^3, N=7 FP64
// gather from 196 to 301 us. The chunks took the N=1, 96^3 FP64 gather from
// 77 us (direct loads past the cap) to 28 us (26 us with a 4096 cap).
#pragma once

#include "kcommon.hpp"

namespace libp_amd {

constexpr int GB_THREADS = 256;
constexpr int GB_ROWS = 2 * GB_THREADS;  // two rows per thread
constexpr int GB_CAP = 2048;             // copies staged per chunk (8 per thread)

__host__ __device__ inline long gather_row_blocks(long Nrows) { return (Nrows + GB_ROWS - 1) / GB_ROWS; }
// grid of a gather that gives each workgroup one row block (a workgroup loops
// over row blocks past 2^20 of them)
inline unsigned row_block_grid(long Nrows) {
  const long nb = gather_row_blocks(Nrows);
  return (unsigned)(nb < 1 ? 1 : (nb > (1L << 20) ? (1L << 20) : nb));
}

// Rows of row block rb handled by this thread: row[u] = rb*GB_ROWS + t + u*GB_THREADS
// (valid where in[u]), their sums in val[u]. s_v holds GB_CAP values. colIds null: the
// copies of a row are v[a, a + len) (slot layout). Every thread of the workgroup calls it
// for the same row block (it synchronises the workgroup).
template <typename T>
struct GatherRows {
  long row[2];
  bool in[2];
  int len[2];  // copies of the row
  T val[2];
};

// load(c) gives the value of the copy at local index c (a plain gather: v[c]).
template <typename T, class Load>
__device__ inline void gather_block_sums_by(const long rb, const int Nrows, const int* __restrict__ rowStarts,
                                            const int* __restrict__ colIds, Load load, T* s_v, GatherRows<T>& g) {
  const int t = threadIdx.x;
  const long r0 = rb * GB_ROWS;
  const long r1 = min(r0 + GB_ROWS, (long)Nrows);
  // the range of copies of the block and of this thread's rows (loads in flight together)
  const int j0 = rowStarts[r0], j1 = rowStarts[r1];
  int a[2], b[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    g.row[u] = r0 + t + u * GB_THREADS;
    g.in[u] = g.row[u] < r1;
    const long rc = min(g.row[u], r1 - 1);
    a[u] = rowStarts[rc];
    b[u] = rowStarts[rc + 1];
    g.len[u] = b[u] - a[u];
  }
  const int n = __builtin_amdgcn_readfirstlane(j1 - j0);
  g.val[0] = g.val[1] = T(0);
  // chunks of GB_CAP copies; n is a scalar, so the loop and the branches around the
  // loads are uniform and no load waits for another
  for (int c0 = 0; c0 < n; c0 += GB_CAP) {
    const int lo = j0 + c0;
    const int m = min(n - c0, GB_CAP);
    const int nl = (m + GB_THREADS - 1) / GB_THREADS;
    int c[GB_CAP / GB_THREADS];
    T x[GB_CAP / GB_THREADS];
#pragma unroll
    for (int u = 0; u < GB_CAP / GB_THREADS; ++u) {
      const int j = min(lo + t + u * GB_THREADS, j1 - 1);
      c[u] = (u < nl) ? (colIds ? colIds[j] : j) : 0;
    }
#pragma unroll
    for (int u = 0; u < GB_CAP / GB_THREADS; ++u) x[u] = (u < nl) ? load(c[u]) : T(0);
    __syncthreads();  // the previous chunk's or row block's sums have read s_v
#pragma unroll
    for (int u = 0; u < GB_CAP / GB_THREADS; ++u)
      if (u < nl && t + u * GB_THREADS < m) s_v[t + u * GB_THREADS] = x[u];
    __syncthreads();
    const int hi = lo + m;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      T s = g.val[u];
      for (int k = max(a[u], lo); k < min(b[u], hi); ++k) s += s_v[k - lo];
      g.val[u] = s;
    }
  }
}

template <typename T>
__device__ inline void gather_block_sums(const long rb, const int Nrows, const int* __restrict__ rowStarts,
                                         const int* __restrict__ colIds, const T* __restrict__ v, T* s_v,
                                         GatherRows<T>& g) {
  gather_block_sums_by(rb, Nrows, rowStarts, colIds, [v](int c) { return v[c]; }, s_v, g);
}

}  // namespace libp_amd
