// Single-rank gather-scatter kernels (CSR rows, one thread per gathered row).
//
// The reference stages blocks of <= 512 nonzeros through shared memory
// (libs/ogs/okl/ogsKernels.okl:86-164). On the C0 hex path every row holds
// 1-8 entries (the copies of one node), so a thread per row needs no LDS
// staging and no barrier. A row of up to 8 entries issues all of its column
// loads, then all of its value loads, before summing: the copies of a node
// lie in different elements, and a loop of dependent loads would leave the
// memory system idle. Sums still run from zero in ascending column order,
// which is the reference's order, so Gather results are bit-identical.
#include "kcommon.hpp"

namespace libp_amd {

// sum_{j in [start,end)} v[colIds[j]] in ascending j; loads issued up front for rows of <= 8 entries
template <typename T>
__device__ inline T gather_row(const int start, const int end, const int* __restrict__ colIds,
                               const T* __restrict__ v) {
  const int len = end - start;
  T val = 0;
  if (len <= 8) {
    int c[8];
    T x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = (k < len) ? colIds[start + k] : 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = (k < len) ? v[c[k]] : T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (k < len) val += x[k];
  } else {
    for (int g = start; g < end; ++g) val += v[colIds[g]];
  }
  return val;
}

template <typename T>
__global__ void __launch_bounds__(256) ogs_gather_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                         const int* __restrict__ colIds, const T* __restrict__ v,
                                                         T* __restrict__ gv) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < Nrows; row += gridDim.x * blockDim.x)
    gv[row] = gather_row<T>(rowStarts[row], rowStarts[row + 1], colIds, v);
}

template <typename T>
__global__ void __launch_bounds__(256) ogs_scatter_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                          const int* __restrict__ colIds, const T* __restrict__ gv,
                                                          T* __restrict__ v) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < Nrows; row += gridDim.x * blockDim.x) {
    const int start = rowStarts[row], end = rowStarts[row + 1];
    const T val = gv[row];
    for (int g = start; g < end; ++g) v[colIds[g]] = val;
  }
}

template <typename T>
static int gather(int Nrows, const int* rs, const int* ci, const T* v, T* gv, void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rs || !ci || !v || !gv) return arg_error("ogs_gather", "null pointer");
  hipLaunchKernelGGL(ogs_gather_kernel<T>, dim3(stream_blocks(Nrows, 256)), dim3(256), 0, as_stream(stream), Nrows,
                     rs, ci, v, gv);
  return launch_status("ogs_gather");
}

template <typename T>
static int scatter(int Nrows, const int* rs, const int* ci, const T* gv, T* v, void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rs || !ci || !v || !gv) return arg_error("ogs_scatter", "null pointer");
  hipLaunchKernelGGL(ogs_scatter_kernel<T>, dim3(stream_blocks(Nrows, 256)), dim3(256), 0, as_stream(stream),
                     Nrows, rs, ci, gv, v);
  return launch_status("ogs_scatter");
}

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_ogs_gather_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v, double* gv,
                                   void* stream) {
  return gather<double>(Nrows, rowStarts, colIds, v, gv, stream);
}
extern "C" int libp_ogs_gather_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v, float* gv,
                                   void* stream) {
  return gather<float>(Nrows, rowStarts, colIds, v, gv, stream);
}
extern "C" int libp_ogs_scatter_f64(int Nrows, const int* rowStarts, const int* colIds, const double* gv, double* v,
                                    void* stream) {
  return scatter<double>(Nrows, rowStarts, colIds, gv, v, stream);
}
extern "C" int libp_ogs_scatter_i32(int Nrows, const int* rowStarts, const int* colIds, const int* gv, int* v,
                                    void* stream) {
  return scatter<int>(Nrows, rowStarts, colIds, gv, v, stream);
}
