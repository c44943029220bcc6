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

// eight lanes per element, lane n = i + 2 j + 4 k holds node n; the tensor
// contractions exchange values with the lanes that differ in one index bit
__global__ void __launch_bounds__(256) coarse_ax_hex1_kernel(const int E, const int* __restrict__ G2L,
                                                             const double* __restrict__ wJ,
                                                             const double* __restrict__ ggeo,
                                                             const double* __restrict__ D, const double lambda,
                                                             const int first, const double* __restrict__ z,
                                                             const double* __restrict__ p,
                                                             const double* __restrict__ s, double* __restrict__ AqL) {
  if (s[8] != 0.0) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = t >> 3, n = t & 7;
  const bool live = e < E;  // every lane takes part in the shuffles
  const int i = n & 1, j = (n >> 1) & 1, k = n >> 2;
  const int lane = threadIdx.x & 63, base = lane & ~7;
  const double beta = first ? 0.0 : s[3];
  double q = 0.0, G00 = 0, G01 = 0, G02 = 0, G11 = 0, G12 = 0, G22 = 0, W = 0;
  if (live) {
    const int g = G2L[(size_t)e * 8 + n];
    q = (g >= 0) ? z[g] + beta * p[g] : 0.0;
    const double* G = ggeo + (size_t)e * 48 + n;
    G00 = G[0];
    G01 = G[8];
    G02 = G[16];
    G11 = G[24];
    G12 = G[32];
    G22 = G[40];
    W = wJ[(size_t)e * 8 + n];
  }
  const double D00 = D[0], D01 = D[1], D10 = D[2], D11 = D[3];
  auto Dm = [&](int row, int col) { return row ? (col ? D11 : D10) : (col ? D01 : D00); };
  // partner values: q at index bit b = 0 and 1
  auto pair = [&](double v, int bit, double& v0, double& v1) {
    const double a0 = __shfl(v, base + (n & ~bit), 64);
    const double a1 = __shfl(v, base + (n | bit), 64);
    v0 = a0;
    v1 = a1;
  };
  double q0, q1;
  pair(q, 1, q0, q1);
  const double qr = Dm(i, 0) * q0 + Dm(i, 1) * q1;
  pair(q, 2, q0, q1);
  const double qs = Dm(j, 0) * q0 + Dm(j, 1) * q1;
  pair(q, 4, q0, q1);
  const double qt = Dm(k, 0) * q0 + Dm(k, 1) * q1;
  const double Gr = G00 * qr + G01 * qs + G02 * qt;
  const double Gs = G01 * qr + G11 * qs + G12 * qt;
  const double Gt = G02 * qr + G12 * qs + G22 * qt;
  double a = lambda * W * q;
  double v0, v1;
  pair(Gr, 1, v0, v1);
  a += Dm(0, i) * v0 + Dm(1, i) * v1;
  pair(Gs, 2, v0, v1);
  a += Dm(0, j) * v0 + Dm(1, j) * v1;
  pair(Gt, 4, v0, v1);
  a += Dm(0, k) * v0 + Dm(1, k) * v1;
  if (live) AqL[(size_t)e * 8 + n] = a;
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
  hipLaunchKernelGGL(coarse_ax_hex1_kernel, dim3((8 * (long)Nelements + 255) / 256), dim3(256), 0, st, Nelements,
                     GlobalToLocal, wJ, ggeo, D, lambda, first, z, p, scalars, AqL);
  // one row per thread: these vectors are small and the kernels latency-bound
  const unsigned blocks = (unsigned)std::min<long>(((long)Nrows + RBS - 1) / RBS, RMAXB);
  hipLaunchKernelGGL(coarse_gather_pap_kernel, dim3(blocks), dim3(RBS), 0, st, Nrows, rowStarts, colIds, AqL, first, z,
                     p, Ap, scalars, (ReduceWS*)ws);
  hipLaunchKernelGGL(coarse_update_kernel, dim3(blocks), dim3(RBS), 0, st, Nrows, invDiag, p, Ap, x, r, z, scalars,
                     (ReduceWS*)ws);
  return launch_status("pcg_jacobi_hex1_iteration");
}
