// Fused vector kernels of the non-blocking (pipelined) CG solvers, with their
// scalars kept on the device (the reference copies block partials to the host
// and waits on an allreduce per fused update).
//
// NBPCG scalars:  [0] gamma=r.z  [1] previous gamma  [2] alpha  [3] beta  [4] delta=p.s  [5] r.r  [6] z.z
// NBFPCG scalars: [0] gamma=u.r  [1] delta=u.w  [2] eta  [3] alpha  [4] beta  [5] r.r
//
// The kernels that write vectors store fence-free block partials; a one-block
// finisher kernel sums them in block order and updates the scalars (a
// per-block agent-scope fence would write back the dirty L2 of every block).
#include "reduce.hpp"

namespace libp_amd {

// p = z + beta p ; s = Z + beta s ; delta = p.s ; alpha = gamma/delta
// (linearSolverUpdateNBPCG.okl update1NBPCG, linearSolverNBPCG.cpp:118-128)
__global__ void __launch_bounds__(RBS) nbpcg_update1_kernel(const int N, const double* __restrict__ z,
                                                            const double* __restrict__ Z, double* __restrict__ p,
                                                            double* __restrict__ s, double* __restrict__ sc,
                                                            ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  const double beta = sc[3];
  double a0 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const double pn = z[n] + beta * p[n];
    const double sn = Z[n] + beta * s[n];
    a0 += pn * sn;
    p[n] = pn;
    s[n] = sn;
  }
  double v[1] = {a0};
  store_block_partials<1>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) nbpcg_update1_finish_kernel(const int nb, double* __restrict__ sc,
                                                                   const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double out[1];
  sum_block_partials<1>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    sc[4] = out[0];
    sc[2] = sc[0] / out[0];
  }
}

// r -= alpha s ; z -= alpha S ; r.z, z.z, r.r ; gamma history and beta
// (update2NBPCG, linearSolverNBPCG.cpp:130-152). first: alpha = 0 and beta = 0.
__global__ void __launch_bounds__(RBS) nbpcg_update2_kernel(const int N, const int first,
                                                            const double* __restrict__ s,
                                                            const double* __restrict__ S, double* __restrict__ r,
                                                            double* __restrict__ z, double* __restrict__ sc,
                                                            ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  const double alpha = first ? 0.0 : sc[2];
  double a0 = 0, a1 = 0, a2 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const double rn = r[n] - alpha * s[n];
    const double zn = z[n] - alpha * S[n];
    a0 += rn * zn;
    a1 += zn * zn;
    a2 += rn * rn;
    r[n] = rn;
    z[n] = zn;
  }
  double v[3] = {a0, a1, a2};
  store_block_partials<3>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) nbpcg_update2_finish_kernel(const int nb, const int first,
                                                                   double* __restrict__ sc,
                                                                   const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double out[3];
  sum_block_partials<3>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    sc[1] = first ? 0.0 : sc[0];
    sc[0] = out[0];
    sc[6] = out[1];
    sc[5] = out[2];
    sc[3] = first ? 0.0 : out[0] / sc[1];
  }
}

// y += sc[k] x  (the delayed x update of NBPCG with the device alpha)
__global__ void __launch_bounds__(256) axpy_dev_kernel(const int N, const double* __restrict__ a,
                                                       const double* __restrict__ x, double* __restrict__ y) {
  const double alpha = *a;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) y[n] += alpha * x[n];
}

// u.r, u.w, r.r ; eta = delta ; alpha = gamma/eta ; beta = 0
// (update0NBFPCG, linearSolverNBFPCG.cpp:93-113)
__global__ void __launch_bounds__(RBS) nbfpcg_update0_kernel(const int N, const double* __restrict__ u,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ w, double* __restrict__ sc,
                                                             ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double a0 = 0, a1 = 0, a2 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const double un = u[n], rn = r[n];
    a0 += un * rn;
    a1 += un * w[n];
    a2 += rn * rn;
  }
  double v[3] = {a0, a1, a2};
  store_block_partials<3>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) nbfpcg_update0_finish_kernel(const int nb, double* __restrict__ sc,
                                                                    const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double out[3];
  sum_block_partials<3>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    sc[0] = out[0];
    sc[1] = out[1];
    sc[5] = out[2];
    sc[2] = out[1];
    sc[3] = out[0] / out[1];
    sc[4] = 0.0;
  }
}

// x += alpha p ; r -= alpha s ; u -= alpha q ; w -= alpha z ; u.r, u.s, u.w, r.r ;
// gamma, beta = -u.s/eta, delta, then eta = delta - beta^2 eta and alpha = gamma/eta
// (update1NBFPCG, linearSolverNBFPCG.cpp:124-168)
__global__ void __launch_bounds__(RBS) nbfpcg_update1_kernel(const int N, const double* __restrict__ p,
                                                             const double* __restrict__ s,
                                                             const double* __restrict__ q,
                                                             const double* __restrict__ z, double* __restrict__ x,
                                                             double* __restrict__ r, double* __restrict__ u,
                                                             double* __restrict__ w, double* __restrict__ sc,
                                                             ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  const double alpha = sc[3];
  double a0 = 0, a1 = 0, a2 = 0;
  double a3 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const double sn = s[n];
    const double xn = x[n] + alpha * p[n];
    const double rn = r[n] - alpha * sn;
    const double un = u[n] - alpha * q[n];
    const double wn = w[n] - alpha * z[n];
    a0 += un * rn;
    a1 += un * sn;
    a2 += un * wn;
    a3 += rn * rn;
    x[n] = xn;
    r[n] = rn;
    u[n] = un;
    w[n] = wn;
  }
  double v[4] = {a0, a1, a2, a3};
  store_block_partials<4>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) nbfpcg_update1_finish_kernel(const int nb, double* __restrict__ sc,
                                                                    const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double out[4];
  sum_block_partials<4>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    const double eta = sc[2];
    const double beta = -out[1] / eta;
    const double delta = out[2];
    const double etaNew = delta - beta * beta * eta;
    sc[0] = out[0];
    sc[1] = delta;
    sc[4] = beta;
    sc[5] = out[3];
    sc[2] = etaNew;
    sc[3] = out[0] / etaNew;
  }
}

// p = u + beta p ; s = w + beta s ; q = m + beta q ; z = n + beta z  (linearSolverNBFPCG.cpp:155-165)
__global__ void __launch_bounds__(256) nbfpcg_update2_kernel(const int N, const double* __restrict__ u,
                                                             const double* __restrict__ w,
                                                             const double* __restrict__ m,
                                                             const double* __restrict__ nn, double* __restrict__ p,
                                                             double* __restrict__ s, double* __restrict__ q,
                                                             double* __restrict__ z, const double* __restrict__ sc) {
  const double beta = sc[4];
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    p[n] = 1.0 * u[n] + beta * p[n];
    s[n] = 1.0 * w[n] + beta * s[n];
    q[n] = 1.0 * m[n] + beta * q[n];
    z[n] = 1.0 * nn[n] + beta * z[n];
  }
}

}  // namespace libp_amd

using namespace libp_amd;

#define NB_CHECK(cond, what) \
  if (cond) return arg_error(what, "null pointer")

extern "C" int libp_nbpcg_update1_f64(int N, const double* z, const double* Z, double* p, double* s, double* scalars,
                                      void* ws, void* stream) {
  NB_CHECK(!z || !Z || !p || !s || !scalars || !ws, "nbpcg_update1");
  const unsigned nb = reduce_blocks(N);
  hipLaunchKernelGGL(nbpcg_update1_kernel, dim3(nb), dim3(RBS), 0, as_stream(stream), N, z, Z, p, s, scalars,
                     (ReduceWS*)ws);
  hipLaunchKernelGGL(nbpcg_update1_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)nb, scalars,
                     (const ReduceWS*)ws);
  return launch_status("nbpcg_update1");
}
extern "C" int libp_nbpcg_update2_f64(int N, int first, const double* s, const double* S, double* r, double* z,
                                      double* scalars, void* ws, void* stream) {
  NB_CHECK(!s || !S || !r || !z || !scalars || !ws, "nbpcg_update2");
  const unsigned nb = reduce_blocks(N);
  hipLaunchKernelGGL(nbpcg_update2_kernel, dim3(nb), dim3(RBS), 0, as_stream(stream), N, first, s, S, r, z, scalars,
                     (ReduceWS*)ws);
  hipLaunchKernelGGL(nbpcg_update2_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)nb, first, scalars,
                     (const ReduceWS*)ws);
  return launch_status("nbpcg_update2");
}
extern "C" int libp_axpy_dev_f64(int N, const double* alpha, const double* x, double* y, void* stream) {
  NB_CHECK(!alpha || !x || !y, "axpy_dev");
  if (N <= 0) return LIBP_OK;
  hipLaunchKernelGGL(axpy_dev_kernel, dim3(stream_blocks(N, 256)), dim3(256), 0, as_stream(stream), N, alpha, x, y);
  return launch_status("axpy_dev");
}
extern "C" int libp_nbfpcg_update0_f64(int N, const double* u, const double* r, const double* w, double* scalars,
                                       void* ws, void* stream) {
  NB_CHECK(!u || !r || !w || !scalars || !ws, "nbfpcg_update0");
  hipLaunchKernelGGL(nbfpcg_update0_kernel, dim3(reduce_blocks(N)), dim3(RBS), 0, as_stream(stream), N, u, r, w,
                     scalars, (ReduceWS*)ws);
  hipLaunchKernelGGL(nbfpcg_update0_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)reduce_blocks(N),
                     scalars, (const ReduceWS*)ws);
  return launch_status("nbfpcg_update0");
}
extern "C" int libp_nbfpcg_update1_f64(int N, const double* p, const double* s, const double* q, const double* z,
                                       double* x, double* r, double* u, double* w, double* scalars, void* ws,
                                       void* stream) {
  NB_CHECK(!p || !s || !q || !z || !x || !r || !u || !w || !scalars || !ws, "nbfpcg_update1");
  const unsigned nb = reduce_blocks(N);
  hipLaunchKernelGGL(nbfpcg_update1_kernel, dim3(nb), dim3(RBS), 0, as_stream(stream), N, p, s, q, z, x, r, u, w,
                     scalars, (ReduceWS*)ws);
  hipLaunchKernelGGL(nbfpcg_update1_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)nb, scalars,
                     (const ReduceWS*)ws);
  return launch_status("nbfpcg_update1");
}
extern "C" int libp_nbfpcg_update2_f64(int N, const double* u, const double* w, const double* m, const double* n,
                                       double* p, double* s, double* q, double* z, const double* scalars,
                                       void* stream) {
  NB_CHECK(!u || !w || !m || !n || !p || !s || !q || !z || !scalars, "nbfpcg_update2");
  if (N <= 0) return LIBP_OK;
  hipLaunchKernelGGL(nbfpcg_update2_kernel, dim3(stream_blocks(N, 256)), dim3(256), 0, as_stream(stream), N, u, w, m,
                     n, p, s, q, z, scalars);
  return launch_status("nbfpcg_update2");
}
