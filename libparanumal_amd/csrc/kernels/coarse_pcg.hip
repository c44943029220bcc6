// One Jacobi-PCG iteration on the N=1 hex operator in three kernels, for the
// FP64 coarse level of the p-multigrid.
//
// The coarse level (8 nodes per element, ~3e4 rows on a 32^3 box) is
// latency-bound: the six launches of a plain PCG iteration (z and r.z, p
// update, Ax, gather, p.Ap, x/r update) cost more than their traffic. Here:
//   A: AqL = A_e (z + beta p) per element, the p update folded into the read
//   B: Ap = gather(AqL); p = z + beta p; p.Ap -> alpha        (one reduction)
//   C: x += alpha p; r -= alpha Ap; z = invD r; r.r and r.z -> beta, STOP
// The arithmetic of each step is that of linearSolverPCG.cpp:107-171 with
// the Jacobi preconditioner (ellipticPreconJacobi.cpp:42-51) and of the
// partial Ax (ellipticAxHex3D.okl:156-295) at N=1. Every kernel returns at
// once when the STOP flag of the scalar block is set, so iterations can be
// queued in chunks (see libp_pcg_update_f64).
#include <algorithm>

#include "reduce.hpp"

namespace libp_amd {

// thread per element; node n = i + 2 j + 4 k
__global__ void __launch_bounds__(256) coarse_ax_hex1_kernel(const int E, const int* __restrict__ G2L,
                                                             const double* __restrict__ wJ,
                                                             const double* __restrict__ ggeo,
                                                             const double* __restrict__ D, const double lambda,
                                                             const int first, const double* __restrict__ z,
                                                             const double* __restrict__ p,
                                                             const double* __restrict__ s, double* __restrict__ AqL) {
  if (s[8] != 0.0) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const double beta = first ? 0.0 : s[3];
  double d[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) d[a][b] = D[a * 2 + b];
  double q[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int g = G2L[(size_t)e * 8 + n];
    q[n] = (g >= 0) ? z[g] + beta * p[g] : 0.0;
  }
  const double* G = ggeo + (size_t)e * 48;
  double Gr[8], Gs[8], Gt[8];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int n = i + 2 * j + 4 * k;
        double qr = 0, qs = 0, qt = 0;
#pragma unroll
        for (int m = 0; m < 2; ++m) {
          qr += d[i][m] * q[m + 2 * j + 4 * k];
          qs += d[j][m] * q[i + 2 * m + 4 * k];
          qt += d[k][m] * q[i + 2 * j + 4 * m];
        }
        const double G00 = G[0 * 8 + n], G01 = G[1 * 8 + n], G02 = G[2 * 8 + n];
        const double G11 = G[3 * 8 + n], G12 = G[4 * 8 + n], G22 = G[5 * 8 + n];
        Gr[n] = G00 * qr + G01 * qs + G02 * qt;
        Gs[n] = G01 * qr + G11 * qs + G12 * qt;
        Gt[n] = G02 * qr + G12 * qs + G22 * qt;
      }
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int n = i + 2 * j + 4 * k;
        double a = lambda * wJ[(size_t)e * 8 + n] * q[n];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
          a += d[m][i] * Gr[m + 2 * j + 4 * k];
          a += d[m][j] * Gs[i + 2 * m + 4 * k];
          a += d[m][k] * Gt[i + 2 * j + 4 * m];
        }
        AqL[(size_t)e * 8 + n] = a;
      }
}

__global__ void __launch_bounds__(RBS) coarse_gather_pap_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                                const int* __restrict__ colIds,
                                                                const double* __restrict__ AqL, const int first,
                                                                const double* __restrict__ z,
                                                                double* __restrict__ p, double* __restrict__ Ap,
                                                                double* __restrict__ s, ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const double beta = first ? 0.0 : s[3];
  double acc = 0;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < Nrows; g += gridDim.x * blockDim.x) {
    const int a = rowStarts[g], b = rowStarts[g + 1];
    double ap = 0;
    {
      const int len = b - a;
      int c[8];
      double xv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) c[k] = (k < len) ? colIds[a + k] : 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) xv[k] = (k < len) ? AqL[c[k]] : 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < len) ap += xv[k];
      for (int j = a + 8; j < b; ++j) ap += AqL[colIds[j]];
    }
    const double pn = first ? z[g] : z[g] + beta * p[g];
    p[g] = pn;
    Ap[g] = ap;
    acc += pn * ap;
  }
  double v[1] = {acc};
  bool last;
  finish_partials<1>(ws, v, s_w, last);
  if (last) {
    double out[1];
    sum_partials<1>(ws, out);
    if (threadIdx.x == 0) {
      s[4] = out[0];
      s[2] = s[0] / out[0];
    }
  }
}

__global__ void __launch_bounds__(RBS) coarse_update_kernel(const int N, const double* __restrict__ invD,
                                                            const double* __restrict__ p,
                                                            const double* __restrict__ Ap, double* __restrict__ x,
                                                            double* __restrict__ r, double* __restrict__ z,
                                                            double* __restrict__ s, ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const double alpha = s[2];
  double a0 = 0, a1 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    x[n] += alpha * p[n];
    const double rn = r[n] - alpha * Ap[n];
    r[n] = rn;
    const double zn = 1.0 * invD[n] * rn;
    z[n] = zn;
    a0 += rn * rn;
    a1 += rn * zn;
  }
  double v[2] = {a0, a1};
  bool last;
  finish_partials<2>(ws, v, s_w, last);
  if (last) {
    double out[2];
    sum_partials<2>(ws, out);
    if (threadIdx.x == 0) {
      s[5] = out[0];
      s[9] += 1.0;
      if (s[7] > 0.0 && out[0] <= s[7]) s[8] = 1.0;
      s[1] = s[0];
      s[0] = out[1];
      s[3] = out[1] / s[1];
    }
  }
}

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_pcg_jacobi_hex1_iteration_f64(int Nelements, const int* GlobalToLocal, const double* wJ,
                                                  const double* ggeo, const double* D, double lambda, int Nrows,
                                                  const int* rowStarts, const int* colIds, const double* invDiag,
                                                  int first, double* x, double* r, double* z, double* p, double* Ap,
                                                  double* AqL, double* scalars, void* ws, void* stream) {
  if (Nelements <= 0 || Nrows <= 0) return LIBP_OK;
  if (!GlobalToLocal || !wJ || !ggeo || !D || !rowStarts || !colIds || !invDiag || !x || !r || !z || !p || !Ap ||
      !AqL || !scalars || !ws)
    return arg_error("pcg_jacobi_hex1_iteration", "null pointer");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(coarse_ax_hex1_kernel, dim3((Nelements + 255) / 256), dim3(256), 0, st, Nelements, GlobalToLocal,
                     wJ, ggeo, D, lambda, first, z, p, scalars, AqL);
  // one row per thread: these vectors are small and the kernels latency-bound
  const unsigned blocks = (unsigned)std::min<long>(((long)Nrows + RBS - 1) / RBS, RMAXB);
  hipLaunchKernelGGL(coarse_gather_pap_kernel, dim3(blocks), dim3(RBS), 0, st, Nrows, rowStarts, colIds, AqL, first, z,
                     p, Ap, scalars, (ReduceWS*)ws);
  hipLaunchKernelGGL(coarse_update_kernel, dim3(blocks), dim3(RBS), 0, st, Nrows, invDiag, p, Ap, x, r, z, scalars,
                     (ReduceWS*)ws);
  return launch_status("pcg_jacobi_hex1_iteration");
}
