// BLAS-1, deterministic reductions to device scalars, fused PCG/FPCG vector
// kernels and Chebyshev smoother steps, FP64 and FP32.
//
// Reductions use the deterministic two-stage scheme of reduce.hpp.
#include <mutex>

#include "reduce.hpp"

namespace libp_amd {
template <typename... P>
inline bool aligned16(const P*... ptrs) {
  return (((((uintptr_t)ptrs) & 15) == 0) && ...);
}
}  // namespace libp_amd

namespace libp_amd {

/* ------------------------------ reductions ------------------------------ */
template <int KIND>  // 0 x.y, 1 x.x, 2 sum x, 3 w.x.y
__global__ void __launch_bounds__(RBS) reduce_kernel(const int N, const double* __restrict__ w,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double acc = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    if (KIND == 0) acc += x[n] * y[n];
    if (KIND == 1) acc += x[n] * x[n];
    if (KIND == 2) acc += x[n];
    if (KIND == 3) acc += w[n] * x[n] * y[n];
  }
  double v[1] = {acc};
  store_block_partials<1>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) reduce_finish_kernel(const int nb, double* __restrict__ result,
                                                            const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double out[1];
  sum_block_partials<1>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    *result = out[0];
  }
}

template <int KIND>
static int reduce(int N, const double* w, const double* x, const double* y, double* result, void* ws, void* stream) {
  if (!x || !result || !ws) return arg_error("reduce", "null pointer");
  const unsigned nb = reduce_blocks(N);
  hipLaunchKernelGGL(reduce_kernel<KIND>, dim3(nb), dim3(RBS), 0, as_stream(stream), N, w, x, y, (ReduceWS*)ws);
  hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)nb, result,
                     (const ReduceWS*)ws);
  return launch_status("reduce");
}

/* -------------------------------- BLAS-1 -------------------------------- */
template <typename T>
__global__ void __launch_bounds__(256) axpy_kernel(const int N, const T alpha, const T* __restrict__ x, const T beta,
                                                   T* __restrict__ y) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x)
    y[n] = (beta != T(0)) ? alpha * x[n] + beta * y[n] : alpha * x[n];
}
template <typename T>
__global__ void __launch_bounds__(256) zaxpy_kernel(const int N, const T alpha, const T* __restrict__ x, const T beta,
                                                    const T* __restrict__ y, T* __restrict__ z) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x)
    z[n] = alpha * x[n] + beta * y[n];
}
template <typename T>
__global__ void __launch_bounds__(256) amx_kernel(const int N, const T alpha, const T* __restrict__ a,
                                                  T* __restrict__ x) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x)
    x[n] = alpha * a[n] * x[n];
}
template <typename T>
__global__ void __launch_bounds__(256) amxpy_kernel(const int N, const T alpha, const T* __restrict__ a,
                                                    const T* __restrict__ x, const T beta, T* __restrict__ y) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x)
    y[n] = (beta != T(0)) ? alpha * a[n] * x[n] + beta * y[n] : alpha * a[n] * x[n];
}
template <int OP>  // 0 set, 1 add, 2 scale
__global__ void __launch_bounds__(256) scalar_kernel(const int N, const double alpha, double* __restrict__ x) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    if (OP == 0) x[n] = alpha;
    if (OP == 1) x[n] += alpha;
    if (OP == 2) x[n] *= alpha;
  }
}
template <typename S, typename D>
__global__ void __launch_bounds__(256) convert_kernel(const int N, const S* __restrict__ x, D* __restrict__ y) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) y[n] = (D)x[n];
}

#define LIBP_LAUNCH_1D(kern, N, ...)                                                                      \
  do {                                                                                                    \
    if ((N) <= 0) return LIBP_OK;                                                                         \
    hipLaunchKernelGGL(kern, dim3(stream_blocks((N), 256)), dim3(256), 0, as_stream(stream), __VA_ARGS__); \
    return launch_status(#kern);                                                                          \
  } while (0)

/* ---------------------------- PCG fused kernels ---------------------------- */
// scalars: [0] rdotz1 [1] rdotz2 [2] alpha [3] beta [4] pAp [5] rdotr [6] zdotAp [7] TOL

// z = invDiag.*r (if invDiag) ; r.z ; (flexible) z.Ap ; then beta.
template <bool JAC>
__global__ void __launch_bounds__(RBS) pcg_rdotz_kernel(const int N, const double* __restrict__ invDiag,
                                                        const double* __restrict__ r, double* __restrict__ z,
                                                        const double* __restrict__ Ap, const int flexible,
                                                        double* __restrict__ s, ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  double a0 = 0, a1 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const double rn = r[n];
    double zn;
    if (JAC) {
      zn = 1.0 * invDiag[n] * rn;
      z[n] = zn;
    } else {
      zn = z[n];
    }
    a0 += rn * zn;
    if (flexible) a1 += zn * Ap[n];
  }
  double v[2] = {a0, a1};
  store_block_partials<2>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) pcg_rdotz_finish_kernel(const int nb, const int flexible,
                                                               double* __restrict__ s,
                                                               const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  double out[2];
  sum_block_partials<2>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    // linearSolverPCG.cpp:116-124 ; beta of the first iteration is handled by the p update
    const double rdotz1_old = s[0];
    s[1] = rdotz1_old;
    s[0] = out[0];
    s[6] = out[1];
    s[3] = flexible ? (-s[2] * out[1] / rdotz1_old) : (out[0] / rdotz1_old);
  }
}

// The FP32 output of a preconditioner widened to z and, in the same pass, the
// r.z and (flexible) z.Ap of pcg_rdotz_kernel<false>: the same grid and the
// same per-thread order, so the partials are those of the two separate calls.
__global__ void __launch_bounds__(RBS) pcg_rdotz_f32in_kernel(const int N, const double* __restrict__ r,
                                                              const float* __restrict__ z32, double* __restrict__ z,
                                                              const double* __restrict__ Ap, const int flexible,
                                                              const double* __restrict__ s, ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  double a0 = 0, a1 = 0;
  // four grid-stride steps per round with their loads issued together, in the
  // order of a plain grid-stride loop (the partial sums are unchanged)
  const int stride = gridDim.x * blockDim.x;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  for (; (long)n + 3L * stride < N; n += 4 * stride) {
    double rv[4], zv[4], av[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int m = n + u * stride;
      rv[u] = __builtin_nontemporal_load(r + m);
      zv[u] = (double)__builtin_nontemporal_load(z32 + m);
      av[u] = flexible ? __builtin_nontemporal_load(Ap + m) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      z[n + u * stride] = zv[u];
      a0 += rv[u] * zv[u];
      if (flexible) a1 += zv[u] * av[u];
    }
  }
  for (; n < N; n += stride) {
    const double rn = r[n];
    const double zn = (double)z32[n];
    z[n] = zn;
    a0 += rn * zn;
    if (flexible) a1 += zn * Ap[n];
  }
  double v[2] = {a0, a1};
  store_block_partials<2>(ws, v, s_w);
}

// p = z + beta*p  (axpy(N, 1, z, beta, p); beta==0 on the first iteration: p = z)
__global__ void __launch_bounds__(256) pcg_update_p_kernel(const int N, const int first, const double* __restrict__ z,
                                                           double* __restrict__ p, const double* __restrict__ s) {
  if (s[8] != 0.0) return;
  const double beta = first ? 0.0 : s[3];
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x)
    p[n] = (beta != 0) ? 1.0 * z[n] + beta * p[n] : 1.0 * z[n];
}

__global__ void __launch_bounds__(RBS) pcg_pap_kernel(const int N, const double* __restrict__ p,
                                                      const double* __restrict__ Ap, double* __restrict__ s,
                                                      ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  double a0 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) a0 += p[n] * Ap[n];
  double v[1] = {a0};
  store_block_partials<1>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) pcg_pap_finish_kernel(const int nb, double* __restrict__ s,
                                                             const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  double out[1];
  sum_block_partials<1>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    s[4] = out[0];
    s[2] = s[0] / out[0];  // alpha = rdotz1/pAp
  }
}

// x += alpha p ; r -= alpha Ap ; r.r   (linearSolverUpdatePCG.okl:29-75)
__global__ void __launch_bounds__(RBS) pcg_update_kernel(const int N, const double* __restrict__ p,
                                                         const double* __restrict__ Ap, double* __restrict__ x,
                                                         double* __restrict__ r, const double* __restrict__ s,
                                                         ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;  // converged in an earlier iteration of this chunk
  const double alpha = s[2];
  double a0 = 0;
  // four grid-stride steps per round, their loads issued together (streamed
  // once: non-temporal); each thread visits its rows in the same order as a
  // plain grid-stride loop, so the partial sums are unchanged
  const int stride = gridDim.x * blockDim.x;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  for (; (long)n + 3L * stride < N; n += 4 * stride) {
    double rv[4], xv[4], pv[4], av[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int m = n + u * stride;
      rv[u] = __builtin_nontemporal_load(r + m);
      xv[u] = __builtin_nontemporal_load(x + m);
      pv[u] = __builtin_nontemporal_load(p + m);
      av[u] = __builtin_nontemporal_load(Ap + m);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int m = n + u * stride;
      double xn = xv[u];
      double rn = rv[u];
      xn += alpha * pv[u];
      rn -= alpha * av[u];
      a0 += rn * rn;
      __builtin_nontemporal_store(xn, x + m);
      __builtin_nontemporal_store(rn, r + m);
    }
  }
  for (; n < N; n += stride) {
    double rn = r[n];
    x[n] += alpha * p[n];
    rn -= alpha * Ap[n];
    a0 += rn * rn;
    r[n] = rn;
  }
  double v[1] = {a0};
  store_block_partials<1>(ws, v, s_w);
}

// r.r of pcg_update_kernel's partials; counts the iteration and raises STOP
// once r.r <= TOL (TOL > 0 only)
__global__ void __launch_bounds__(RBS) pcg_update_finish_kernel(const int nb, double* __restrict__ s,
                                                                const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  double out[1];
  sum_block_partials<1>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    s[5] = out[0];
    s[9] += 1.0;
    if (s[7] > 0.0 && out[0] <= s[7]) s[8] = 1.0;
  }
}

// x += alpha p; r -= alpha Ap; z = invDiag.*r; r.r and r.z in one pass: the
// update of linearSolverPCG.cpp:142-160 fused with the Jacobi preconditioner
// and the r.z of the next iteration (:107-124). Its finisher writes r.r,
// counts the iteration, raises STOP once r.r <= TOL, shifts r.z (RDOTZ2 <-
// RDOTZ1 <- r.z) and sets beta = r.z / old r.z.
// The same update with 16-byte accesses: each thread takes two pairs of rows per
// step (four rows, all loads issued before any store), so four times the bytes of
// the one-row loop are in flight per wave; the last row of an odd N is done by
// one thread after the loop. Needs 16-byte aligned vectors.
typedef double pcg_v2d __attribute__((ext_vector_type(2)));
// non-temporal 16-byte load and store: each vector is streamed once per pass
// (the 13.7M-row Jacobi update: 185 -> 177 us; 37.6M rows: 518 -> 470 us)
__device__ inline pcg_v2d ntl(const pcg_v2d* a) { return __builtin_nontemporal_load(a); }
__device__ inline void nts(const pcg_v2d v, pcg_v2d* a) { __builtin_nontemporal_store(v, a); }

__device__ inline void jacobi_pair(const pcg_v2d d, const pcg_v2d pv, const pcg_v2d ap, pcg_v2d& xv, pcg_v2d& rv,
                                   pcg_v2d& zv, const double alpha, double& a0, double& a1) {
  xv.x += alpha * pv.x;
  xv.y += alpha * pv.y;
  rv.x -= alpha * ap.x;
  rv.y -= alpha * ap.y;
  zv.x = 1.0 * d.x * rv.x;
  zv.y = 1.0 * d.y * rv.y;
  a0 += rv.x * rv.x;
  a0 += rv.y * rv.y;
  a1 += rv.x * zv.x;
  a1 += rv.y * zv.y;
}

__global__ void __launch_bounds__(RBS) pcg_jacobi_update2_kernel(const int N, const double* __restrict__ invDiag,
                                                                 const double* __restrict__ p,
                                                                 const double* __restrict__ Ap,
                                                                 double* __restrict__ x, double* __restrict__ r,
                                                                 double* __restrict__ z, const double* __restrict__ s,
                                                                 ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const double alpha = s[2];
  const long N2 = N / 2, S = (long)gridDim.x * blockDim.x;
  const pcg_v2d *d2 = (const pcg_v2d*)invDiag, *p2 = (const pcg_v2d*)p, *A2 = (const pcg_v2d*)Ap;
  pcg_v2d *x2 = (pcg_v2d*)x, *r2 = (pcg_v2d*)r, *z2 = (pcg_v2d*)z;
  double a0 = 0, a1 = 0;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (; i + S < N2; i += 2 * S) {
    const pcg_v2d da = ntl(d2 + i), db = ntl(d2 + i + S), pa = ntl(p2 + i), pb = ntl(p2 + i + S);
    const pcg_v2d Aa = ntl(A2 + i), Ab = ntl(A2 + i + S);
    pcg_v2d xa = ntl(x2 + i), xb = ntl(x2 + i + S), ra = ntl(r2 + i), rb = ntl(r2 + i + S), za, zb;
    jacobi_pair(da, pa, Aa, xa, ra, za, alpha, a0, a1);
    jacobi_pair(db, pb, Ab, xb, rb, zb, alpha, a0, a1);
    nts(xa, x2 + i);
    nts(xb, x2 + i + S);
    nts(ra, r2 + i);
    nts(rb, r2 + i + S);
    nts(za, z2 + i);
    nts(zb, z2 + i + S);
  }
  for (; i < N2; i += S) {
    pcg_v2d xa = x2[i], ra = r2[i], za;
    jacobi_pair(d2[i], p2[i], A2[i], xa, ra, za, alpha, a0, a1);
    x2[i] = xa;
    r2[i] = ra;
    z2[i] = za;
  }
  if ((N & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const int n = N - 1;
    double rn = r[n];
    x[n] += alpha * p[n];
    rn -= alpha * Ap[n];
    r[n] = rn;
    const double zn = 1.0 * invDiag[n] * rn;
    z[n] = zn;
    a0 += rn * rn;
    a1 += rn * zn;
  }
  double v[2] = {a0, a1};
  store_block_partials<2>(ws, v, s_w);
}

// p = z + beta p with 16-byte non-temporal accesses, one pair of rows per
// thread over a grid that covers the vector (no grid-stride loop). On the 48^3,
// N=7 vector (37.6M rows) this form runs at 6.6 TB/s against 4.9 TB/s for two
// pairs per thread per step over 8192 blocks with cached accesses (137 against
// 182 us); each change alone gave part of it (full grid 156, non-temporal 168 us).
__global__ void __launch_bounds__(256) pcg_update_p2_kernel(const int N, const int first, const double* __restrict__ z,
                                                            double* __restrict__ p, const double* __restrict__ s) {
  if (s[8] != 0.0) return;
  const double beta = first ? 0.0 : s[3];
  const long N2 = N / 2;
  const pcg_v2d* z2 = (const pcg_v2d*)z;
  pcg_v2d* p2 = (pcg_v2d*)p;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < N2) {
    const pcg_v2d zv = __builtin_nontemporal_load(z2 + i);
    pcg_v2d o;
    if (beta != 0) {
      const pcg_v2d pv = __builtin_nontemporal_load(p2 + i);
      o = 1.0 * zv + beta * pv;
    } else {
      o = 1.0 * zv;
    }
    __builtin_nontemporal_store(o, p2 + i);
  }
  if ((N & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const int n = N - 1;
    p[n] = (beta != 0) ? 1.0 * z[n] + beta * p[n] : 1.0 * z[n];
  }
}

__global__ void __launch_bounds__(RBS) pcg_jacobi_update_kernel(const int N, const double* __restrict__ invDiag,
                                                                const double* __restrict__ p,
                                                                const double* __restrict__ Ap,
                                                                double* __restrict__ x, double* __restrict__ r,
                                                                double* __restrict__ z, const double* __restrict__ s,
                                                                ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const double alpha = s[2];
  double a0 = 0, a1 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    double rn = r[n];
    x[n] += alpha * p[n];
    rn -= alpha * Ap[n];
    r[n] = rn;
    const double zn = 1.0 * invDiag[n] * rn;
    z[n] = zn;
    a0 += rn * rn;
    a1 += rn * zn;
  }
  double v[2] = {a0, a1};
  store_block_partials<2>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) pcg_jacobi_update_finish_kernel(const int nb, double* __restrict__ s,
                                                                       const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  double out[2];
  sum_block_partials<2>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    s[5] = out[0];
    s[9] += 1.0;
    if (s[7] > 0.0 && out[0] <= s[7]) s[8] = 1.0;
    s[1] = s[0];
    s[0] = out[1];
    s[3] = out[1] / s[1];
  }
}

/* ------------------------------ Chebyshev ------------------------------ */
// RES = invDiag.*(r - Ax) [Ax==NULL: invDiag.*r] ; d = invTheta*RES ; X = d (xIsZero) or X += d
template <typename T>
__global__ void __launch_bounds__(256) cheb_start_kernel(const int N, const T* __restrict__ invDiag,
                                                         const T* __restrict__ r, const T* __restrict__ Ax,
                                                         const T invTheta, T* __restrict__ RES, T* __restrict__ d,
                                                         T* __restrict__ X, const int xIsZero) {
  // no a*b+c contraction: the gather-fused form (ogs_gather_cheb_kernel) must round the same way
#pragma clang fp contract(off)
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const T res = Ax ? T(1) * invDiag[n] * (T(1) * r[n] - T(1) * Ax[n]) : T(1) * invDiag[n] * r[n];
    const T dn = invTheta * res;
    RES[n] = res;
    d[n] = dn;
    X[n] = xIsZero ? dn : X[n] + dn;
  }
}
// RES -= invDiag.*Ad ; d = c1*RES + c2*d ; X += d
template <typename T>
__global__ void __launch_bounds__(256) cheb_step_kernel(const int N, const T* __restrict__ invDiag,
                                                        const T* __restrict__ Ad, const T c1, const T c2,
                                                        T* __restrict__ RES, T* __restrict__ d, T* __restrict__ X) {
#pragma clang fp contract(off)
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const T res = T(-1) * invDiag[n] * Ad[n] + RES[n];
    const T dn = c1 * res + c2 * d[n];
    RES[n] = res;
    d[n] = dn;
    X[n] = X[n] + dn;
  }
}

}  // namespace libp_amd

using namespace libp_amd;

// HBM probes (no reference counterpart; measurement only): a read-only sweep
// that sums x into one partial per block, and a copy y = x. A sweep over grid
// sizes, access widths and cache hints on MI355X gave the highest read rate
// (7.27 TB/s over 4 GiB) with 8-byte non-temporal loads, eight in flight per
// thread, on 65536 blocks of 256 threads, and the highest copy rate (5.18 TB/s)
// with 16-byte non-temporal loads and stores on 16384 blocks.
typedef double hbm_v2d __attribute__((ext_vector_type(2)));
constexpr int HBM_READ_BLOCKS = 65536;
constexpr int HBM_COPY_BLOCKS = 16384;

__global__ void __launch_bounds__(256) hbm_read_kernel(const long n, const double* __restrict__ x,
                                                       double* __restrict__ part) {
  __shared__ double s_w[4];
  const long S = (long)gridDim.x * blockDim.x;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double acc = 0;
  for (; i + 7 * S < n; i += 8 * S) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(x + i + u * S);
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; i < n; i += S) acc += __builtin_nontemporal_load(x + i);
  const double b = block_sum(acc, s_w);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ void __launch_bounds__(256) hbm_copy_kernel(const long n2, const hbm_v2d* __restrict__ x,
                                                       hbm_v2d* __restrict__ y) {
  const long S = (long)gridDim.x * blockDim.x;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (; i + 3 * S < n2; i += 4 * S) {
    const hbm_v2d a = __builtin_nontemporal_load(x + i), b = __builtin_nontemporal_load(x + i + S);
    const hbm_v2d c = __builtin_nontemporal_load(x + i + 2 * S), d = __builtin_nontemporal_load(x + i + 3 * S);
    __builtin_nontemporal_store(a, y + i);
    __builtin_nontemporal_store(b, y + i + S);
    __builtin_nontemporal_store(c, y + i + 2 * S);
    __builtin_nontemporal_store(d, y + i + 3 * S);
  }
  for (; i < n2; i += S) __builtin_nontemporal_store(__builtin_nontemporal_load(x + i), y + i);
}

extern "C" long libp_reduce_ws_bytes(void) { return (long)sizeof(ReduceWS); }

extern "C" int libp_inner_prod_f64(int N, const double* x, const double* y, double* result, void* ws, void* stream) {
  return reduce<0>(N, nullptr, x, y, result, ws, stream);
}
extern "C" int libp_weighted_inner_prod_f64(int N, const double* w, const double* x, const double* y, double* result,
                                            void* ws, void* stream) {
  return reduce<3>(N, w, x, y, result, ws, stream);
}
extern "C" int libp_norm2sq_f64(int N, const double* x, double* result, void* ws, void* stream) {
  return reduce<1>(N, nullptr, x, nullptr, result, ws, stream);
}
extern "C" int libp_sum_f64(int N, const double* x, double* result, void* ws, void* stream) {
  return reduce<2>(N, nullptr, x, nullptr, result, ws, stream);
}

extern "C" int libp_axpy_f64(int N, double alpha, const double* x, double beta, double* y, void* stream) {
  LIBP_LAUNCH_1D(axpy_kernel<double>, N, N, alpha, x, beta, y);
}
extern "C" int libp_axpy_f32(int N, float alpha, const float* x, float beta, float* y, void* stream) {
  LIBP_LAUNCH_1D(axpy_kernel<float>, N, N, alpha, x, beta, y);
}
extern "C" int libp_zaxpy_f64(int N, double alpha, const double* x, double beta, const double* y, double* z,
                              void* stream) {
  LIBP_LAUNCH_1D(zaxpy_kernel<double>, N, N, alpha, x, beta, y, z);
}
extern "C" int libp_amx_f64(int N, double alpha, const double* a, double* x, void* stream) {
  LIBP_LAUNCH_1D(amx_kernel<double>, N, N, alpha, a, x);
}
extern "C" int libp_amx_f32(int N, float alpha, const float* a, float* x, void* stream) {
  LIBP_LAUNCH_1D(amx_kernel<float>, N, N, alpha, a, x);
}
extern "C" int libp_amxpy_f64(int N, double alpha, const double* a, const double* x, double beta, double* y,
                              void* stream) {
  LIBP_LAUNCH_1D(amxpy_kernel<double>, N, N, alpha, a, x, beta, y);
}
extern "C" int libp_amxpy_f32(int N, float alpha, const float* a, const float* x, float beta, float* y, void* stream) {
  LIBP_LAUNCH_1D(amxpy_kernel<float>, N, N, alpha, a, x, beta, y);
}
// n device scalars into pinned host memory mapped for the device (zero-copy):
// a one-block kernel instead of a runtime device-to-host copy. With an event
// behind it, a host read of the PCG scalars costs ~13 us of idle GPU against
// ~25 us for a runtime copy and a stream synchronisation.
__global__ void __launch_bounds__(64) scalars_to_host_kernel(const int n, const double* __restrict__ src,
                                                             double* __restrict__ dst) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = src[i];
}
extern "C" int libp_scalars_to_host_f64(int n, const double* src, double* hostDst, void* stream) {
  if (n < 0 || n > 64) return arg_error("scalars_to_host", "n must be in [0,64]");
  if (n == 0) return LIBP_OK;
  if (!src || !hostDst) return arg_error("scalars_to_host", "null pointer");
  hipLaunchKernelGGL(scalars_to_host_kernel, dim3(1), dim3(64), 0, as_stream(stream), n, src, hostDst);
  return launch_status("scalars_to_host");
}
extern "C" int libp_set_f64(int N, double alpha, double* x, void* stream) {
  LIBP_LAUNCH_1D(scalar_kernel<0>, N, N, alpha, x);
}
extern "C" int libp_add_f64(int N, double alpha, double* x, void* stream) {
  LIBP_LAUNCH_1D(scalar_kernel<1>, N, N, alpha, x);
}
extern "C" int libp_scale_f64(int N, double alpha, double* x, void* stream) {
  LIBP_LAUNCH_1D(scalar_kernel<2>, N, N, alpha, x);
}
extern "C" int libp_convert_f64_to_f32(int N, const double* x, float* y, void* stream) {
  LIBP_LAUNCH_1D((convert_kernel<double, float>), N, N, x, y);
}
extern "C" int libp_convert_f32_to_f64(int N, const float* x, double* y, void* stream) {
  LIBP_LAUNCH_1D((convert_kernel<float, double>), N, N, x, y);
}

extern "C" int libp_pcg_rdotz_f64(int N, const double* r, const double* z, const double* Ap, int flexible,
                                  double* scalars, void* ws, void* stream) {
  hipLaunchKernelGGL(pcg_rdotz_kernel<false>, dim3(reduce_blocks(N)), dim3(RBS), 0, as_stream(stream), N, nullptr,
                     r, const_cast<double*>(z), Ap, flexible, scalars, (ReduceWS*)ws);
  hipLaunchKernelGGL(pcg_rdotz_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)reduce_blocks(N), flexible, scalars,
                     (const ReduceWS*)ws);
  return launch_status("pcg_rdotz");
}
extern "C" int libp_pcg_rdotz_from_f32_f64(int N, const double* r, const float* z32, double* z, const double* Ap,
                                           int flexible, double* scalars, void* ws, void* stream) {
  if (!r || !z32 || !z || !scalars || !ws || (flexible && !Ap)) return arg_error("pcg_rdotz_from_f32", "null pointer");
  const unsigned nb = reduce_blocks(N);
  hipLaunchKernelGGL(pcg_rdotz_f32in_kernel, dim3(nb), dim3(RBS), 0, as_stream(stream), N, r, z32, z, Ap, flexible,
                     scalars, (ReduceWS*)ws);
  hipLaunchKernelGGL(pcg_rdotz_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)nb, flexible, scalars,
                     (const ReduceWS*)ws);
  return launch_status("pcg_rdotz_from_f32");
}
extern "C" int libp_pcg_jacobi_rdotz_f64(int N, const double* invDiag, const double* r, double* z, const double* Ap,
                                         int flexible, double* scalars, void* ws, void* stream) {
  hipLaunchKernelGGL(pcg_rdotz_kernel<true>, dim3(reduce_blocks(N)), dim3(RBS), 0, as_stream(stream), N, invDiag, r,
                     z, Ap, flexible, scalars, (ReduceWS*)ws);
  hipLaunchKernelGGL(pcg_rdotz_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)reduce_blocks(N), flexible, scalars,
                     (const ReduceWS*)ws);
  return launch_status("pcg_jacobi_rdotz");
}
// relative stopping threshold formed on the device (no host read of r0.r0)
__global__ void pcg_relative_tol_kernel(const double tol2, double* __restrict__ s) {
  if (threadIdx.x == 0) {
    const double t = tol2 * s[5];
    s[7] = (t > 0.0) ? t : 1e-300;
    if (s[5] == 0.0) s[8] = 1.0;
  }
}
extern "C" int libp_pcg_relative_tol_f64(double tol, double* scalars, void* stream) {
  if (!scalars) return arg_error("pcg_relative_tol", "null pointer");
  hipLaunchKernelGGL(pcg_relative_tol_kernel, dim3(1), dim3(64), 0, as_stream(stream), tol * tol, scalars);
  return launch_status("pcg_relative_tol");
}

static void pcg_jacobi_update_launch(int N, const double* invDiag, const double* p, const double* Ap, double* x,
                                     double* r, double* z, double* scalars, ReduceWS* ws, hipStream_t st) {
  const unsigned nb = reduce_blocks(N);
  if (aligned16(invDiag, p, Ap, x, r, z))
    hipLaunchKernelGGL(pcg_jacobi_update2_kernel, dim3(nb), dim3(RBS), 0, st, N, invDiag, p, Ap, x, r, z, scalars,
                       ws);
  else
    hipLaunchKernelGGL(pcg_jacobi_update_kernel, dim3(nb), dim3(RBS), 0, st, N, invDiag, p, Ap, x, r, z, scalars, ws);
  hipLaunchKernelGGL(pcg_jacobi_update_finish_kernel, dim3(1), dim3(RBS), 0, st, (int)nb, scalars,
                     (const ReduceWS*)ws);
}

extern "C" int libp_pcg_jacobi_update_f64(int N, const double* invDiag, const double* p, const double* Ap,
                                          double* x, double* r, double* z, double* scalars, void* ws, void* stream) {
  if (!invDiag || !p || !Ap || !x || !r || !z || !scalars || !ws) return arg_error("pcg_jacobi_update", "null pointer");
  if (N < 0) return arg_error("pcg_jacobi_update", "N < 0");
  pcg_jacobi_update_launch(N, invDiag, p, Ap, x, r, z, scalars, (ReduceWS*)ws, as_stream(stream));
  return launch_status("pcg_jacobi_update");
}

// Scratch of libp_pcg_jacobi_update_noalloc_f64: the block partials of the
// fused update in a fixed device array instead of a caller workspace. Every
// GPU the library runs on holds its own copy (32 KiB) until the process ends.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): each device's address comes from
// hipGetSymbolAddress, once per device with that device current.
static __device__ __align__(256) ReduceWS scratch;

namespace {
constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mutex;
ReduceWS* g_scratch_ptr[kScratchMaxDevices] = {};

int scratch_for_current_device(ReduceWS** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_last_error("pcg_jacobi_update_noalloc", hipGetErrorString(e));
    return LIBP_ERR_HIP;
  }
  if (dev < 0 || dev >= kScratchMaxDevices) return arg_error("pcg_jacobi_update_noalloc", "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_last_error("pcg_jacobi_update_noalloc: hipGetSymbolAddress", hipGetErrorString(e));
      return LIBP_ERR_HIP;
    }
    g_scratch_ptr[dev] = (ReduceWS*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return LIBP_OK;
}
}  // namespace

extern "C" long libp_pcg_jacobi_update_noalloc_max_n(void) { return LIBP_NOALLOC_MAX_N; }

extern "C" int libp_pcg_jacobi_update_noalloc_f64(int N, const double* invDiag, const double* p, const double* Ap,
                                                  double* x, double* r, double* z, double* scalars, void* stream) {
  if (!invDiag || !p || !Ap || !x || !r || !z || !scalars)
    return arg_error("pcg_jacobi_update_noalloc", "null pointer");
  if (N < 0 || (long)N > LIBP_NOALLOC_MAX_N)
    return arg_error("pcg_jacobi_update_noalloc", "N outside [0, LIBP_NOALLOC_MAX_N]");
  if ((long)reduce_blocks(N) * (long)sizeof(double) > (long)sizeof(scratch.part[0]))
    return arg_error("pcg_jacobi_update_noalloc", "block partials exceed the static scratch");
  ReduceWS* ws = nullptr;
  const int st = scratch_for_current_device(&ws);
  if (st != LIBP_OK) return st;
  pcg_jacobi_update_launch(N, invDiag, p, Ap, x, r, z, scalars, ws, as_stream(stream));
  return launch_status("pcg_jacobi_update_noalloc");
}
extern "C" int libp_pcg_update_p_f64(int N, int first_iter, const double* z, double* p, const double* scalars,
                                     void* stream) {
  if (N > 0 && z && p && scalars && aligned16(z, p)) {
    hipLaunchKernelGGL(pcg_update_p2_kernel, dim3(stream_blocks((N + 1) / 2, 256, 1L << 30)), dim3(256), 0,
                       as_stream(stream), N, first_iter, z, p, scalars);
    return launch_status("pcg_update_p");
  }
  LIBP_LAUNCH_1D(pcg_update_p_kernel, N, N, first_iter, z, p, scalars);
}
extern "C" int libp_pcg_pap_f64(int N, const double* p, const double* Ap, double* scalars, void* ws, void* stream) {
  hipLaunchKernelGGL(pcg_pap_kernel, dim3(reduce_blocks(N)), dim3(RBS), 0, as_stream(stream), N, p, Ap, scalars,
                     (ReduceWS*)ws);
  hipLaunchKernelGGL(pcg_pap_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)reduce_blocks(N), scalars,
                     (const ReduceWS*)ws);
  return launch_status("pcg_pap");
}
extern "C" int libp_pcg_update_f64(int N, const double* p, const double* Ap, double* x, double* r, double* scalars,
                                   void* ws, void* stream) {
  const unsigned nb = reduce_blocks(N);
  hipLaunchKernelGGL(pcg_update_kernel, dim3(nb), dim3(RBS), 0, as_stream(stream), N, p, Ap, x, r, scalars,
                     (ReduceWS*)ws);
  hipLaunchKernelGGL(pcg_update_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)nb, scalars,
                     (const ReduceWS*)ws);
  return launch_status("pcg_update");
}

extern "C" int libp_cheb_start_f32(int N, const float* invDiag, const float* r, const float* Ax, float invTheta,
                                   float* RES, float* d, float* X, int xIsZero, void* stream) {
  LIBP_LAUNCH_1D(cheb_start_kernel<float>, N, N, invDiag, r, Ax, invTheta, RES, d, X, xIsZero);
}
extern "C" int libp_cheb_step_f32(int N, const float* invDiag, const float* Ad, float c1, float c2, float* RES,
                                  float* d, float* X, void* stream) {
  LIBP_LAUNCH_1D(cheb_step_kernel<float>, N, N, invDiag, Ad, c1, c2, RES, d, X);
}
extern "C" int libp_cheb_start_f64(int N, const double* invDiag, const double* r, const double* Ax, double invTheta,
                                   double* RES, double* d, double* X, int xIsZero, void* stream) {
  LIBP_LAUNCH_1D(cheb_start_kernel<double>, N, N, invDiag, r, Ax, invTheta, RES, d, X, xIsZero);
}
extern "C" int libp_cheb_step_f64(int N, const double* invDiag, const double* Ad, double c1, double c2, double* RES,
                                  double* d, double* X, void* stream) {
  LIBP_LAUNCH_1D(cheb_step_kernel<double>, N, N, invDiag, Ad, c1, c2, RES, d, X);
}

extern "C" int libp_hbm_probe_f64(int mode, long n, const double* x, double* y, void* stream) {
  if (n <= 0) return LIBP_OK;
  if (!x || !y || (n & 1) || ((uintptr_t)x & 15) || ((uintptr_t)y & 15))
    return arg_error("hbm_probe", "null or unaligned pointer, or odd length");
  if (mode == 0) {
    if (n < HBM_READ_BLOCKS) return arg_error("hbm_probe", "read mode needs n >= 65536 (y holds the block sums)");
    hipLaunchKernelGGL(hbm_read_kernel, dim3(HBM_READ_BLOCKS), dim3(256), 0, as_stream(stream), n, x, y);
  } else if (mode == 1) {
    hipLaunchKernelGGL(hbm_copy_kernel, dim3(HBM_COPY_BLOCKS), dim3(256), 0, as_stream(stream), n / 2,
                       (const hbm_v2d*)x, (hbm_v2d*)y);
  } else {
    return arg_error("hbm_probe", "mode must be 0 (read) or 1 (copy)");
  }
  return launch_status("hbm_probe");
}
