// Sparse and dense kernels of the algebraic multigrid that solves the p-MG
// base level (the reference's parAlmond path: okl/SpMVcsr.okl, dGEMV.okl,
// SmoothChebyshev.okl). Matrices are CSR in FP64.
//
// csr_spmv: y = alpha A x + beta z (z may be y, or null for beta = 0), a
// group of LANES lanes per row; each lane
// sums a strided slice of the row in order, then the group reduces with
// shuffles in a fixed tree, so results are reproducible run to run.
// dense_gemv: y = M x for the dense inverse of the coarsest level, one
// wavefront per row.
#include "kcommon.hpp"

namespace libp_amd {

template <int LANES>
__global__ void __launch_bounds__(256) csr_spmv_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                       const int* __restrict__ cols, const double* __restrict__ vals,
                                                       const double alpha, const double* __restrict__ x,
                                                       const double beta, const double* z, double* y) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int sub = threadIdx.x & (LANES - 1);
  const long row = t / LANES;
  // every lane of a group takes part in the shuffles, rows past the end too
  double acc = 0.0;
  int a = 0, b = 0;
  if (row < Nrows) {
    a = rowStarts[row];
    b = rowStarts[row + 1];
  }
  for (int j = a + sub; j < b; j += LANES) acc += vals[j] * x[cols[j]];
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
  if (row < Nrows && sub == 0) {
    double v = alpha * acc;
    if (z) v += beta * z[row];
    y[row] = v;
  }
}

__global__ void __launch_bounds__(256) dense_gemv_kernel(const int n, const double* __restrict__ M,
                                                         const double* __restrict__ x, double* __restrict__ y) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  double acc = 0.0;
  if (row < n)
    for (int j = l; j < n; j += 64) acc += M[(long)row * n + j] * x[j];
  acc = wave_sum(acc);
  if (row < n && l == 0) y[row] = acc;
}

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_csr_spmv_f64(int Nrows, const int* rowStarts, const int* cols, const double* vals, double alpha,
                                 const double* x, double beta, const double* z, double* y, int lanesPerRow,
                                 void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rowStarts || !cols || !vals || !x || !y) return arg_error("csr_spmv", "null pointer");
  hipStream_t st = as_stream(stream);
  const long threads = (long)Nrows * lanesPerRow;
  const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
  switch (lanesPerRow) {
    case 1:
      hipLaunchKernelGGL(csr_spmv_kernel<1>, grid, block, 0, st, Nrows, rowStarts, cols, vals, alpha, x, beta, z, y);
      break;
    case 8:
      hipLaunchKernelGGL(csr_spmv_kernel<8>, grid, block, 0, st, Nrows, rowStarts, cols, vals, alpha, x, beta, z, y);
      break;
    case 64:
      hipLaunchKernelGGL(csr_spmv_kernel<64>, grid, block, 0, st, Nrows, rowStarts, cols, vals, alpha, x, beta, z, y);
      break;
    default:
      return arg_error("csr_spmv", "lanesPerRow must be 1, 8 or 64");
  }
  return launch_status("csr_spmv");
}

extern "C" int libp_dense_gemv_f64(int n, const double* M, const double* x, double* y, void* stream) {
  if (n <= 0) return LIBP_OK;
  if (!M || !x || !y) return arg_error("dense_gemv", "null pointer");
  hipLaunchKernelGGL(dense_gemv_kernel, dim3((n + 3) / 4), dim3(256), 0, as_stream(stream), n, M, x, y);
  return launch_status("dense_gemv");
}
