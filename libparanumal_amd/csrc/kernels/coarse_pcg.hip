// One Jacobi-PCG iteration on the N=1 hex operator in three kernels, for the
// FP64 coarse level of the p-multigrid.
//
// The coarse level (8 nodes per element, ~3e4 rows on a 32^3 box) is
// latency-bound: the six launches of a plain PCG iteration (z and r.z, p
// update, Ax, gather, p.Ap, x/r update) cost more than their traffic. Here:
//   A: r.r, r.z <- partials of C; stop test; beta;
//      AqL = A_e (z + beta p) per element, the p update folded into the read
//   B: Ap = gather(AqL); p = z + beta p; partials of p.Ap
//   C: p.Ap <- partials of B; alpha; x += alpha p; r -= alpha Ap; z = invD r;
//      partials of r.r and r.z
// A reduction is finished by the kernel that needs it: every block of that
// kernel sums all partials of the previous kernel in the same fixed order, so
// all blocks hold bit-identical scalars, and the kernel boundary makes the
// partials visible. No block fences, no tickets: on this GPU an agent-scope
// fence writes back the L2, which in these small kernels cost more than the
// work. Block 0 publishes the scalars for the host. The arithmetic of each
// step is that of linearSolverPCG.cpp:107-171 with the Jacobi preconditioner
// (ellipticPreconJacobi.cpp:42-51) and of the partial Ax
// (ellipticAxHex3D.okl:156-295) at N=1. Every kernel returns at once when
// the STOP flag of the scalar block is set, so iterations can be queued in
// chunks; coarse_check_kernel applies the stopping test to the last queued
// iteration before the host reads the flag.
#include <algorithm>

#include "reduce.hpp"

namespace libp_amd {

// eight lanes per element, lane n = i + 2 j + 4 k holds node n; the tensor
// contractions exchange values with the lanes that differ in one index bit.
// Every lane of the wavefront must call this (the shuffles read all lanes);
// q and the factors are zero on lanes past the last element.
struct Hex1Node {
  double q, G00, G01, G02, G11, G12, G22, W;
};

__device__ inline Hex1Node hex1_load(const int e, const int n, const int* __restrict__ G2L,
                                     const double* __restrict__ wJ, const double* __restrict__ ggeo,
                                     const double* __restrict__ z, const double* __restrict__ p, const double beta,
                                     const bool withP) {
  Hex1Node v;
  const int g = G2L[(size_t)e * 8 + n];
  const int gc = max(g, 0);  // unconditional loads: a masked node reads entry 0 and takes 0
  const double zv = z[gc];
  const double pv = withP ? p[gc] : 0.0;
  v.q = (g < 0) ? 0.0 : (withP ? zv + beta * pv : zv);
  const double* G = ggeo + (size_t)e * 48 + n;
  v.G00 = G[0];
  v.G01 = G[8];
  v.G02 = G[16];
  v.G11 = G[24];
  v.G12 = G[32];
  v.G22 = G[40];
  v.W = wJ[(size_t)e * 8 + n];
  return v;
}

__device__ inline double hex1_apply(const Hex1Node& v, const int n, const double* __restrict__ D,
                                    const double lambda) {
  const int i = n & 1, j = (n >> 1) & 1, k = n >> 2;
  const int lane = threadIdx.x & 63, base = lane & ~7;
  const double D00 = D[0], D01 = D[1], D10 = D[2], D11 = D[3];
  auto Dm = [&](int row, int col) { return row ? (col ? D11 : D10) : (col ? D01 : D00); };
  // partner values: v at index bit b = 0 and 1
  auto pair = [&](double x, int bit, double& x0, double& x1) {
    x0 = __shfl(x, base + (n & ~bit), 64);
    x1 = __shfl(x, base + (n | bit), 64);
  };
  double q0, q1;
  pair(v.q, 1, q0, q1);
  const double qr = Dm(i, 0) * q0 + Dm(i, 1) * q1;
  pair(v.q, 2, q0, q1);
  const double qs = Dm(j, 0) * q0 + Dm(j, 1) * q1;
  pair(v.q, 4, q0, q1);
  const double qt = Dm(k, 0) * q0 + Dm(k, 1) * q1;
  const double Gr = v.G00 * qr + v.G01 * qs + v.G02 * qt;
  const double Gs = v.G01 * qr + v.G11 * qs + v.G12 * qt;
  const double Gt = v.G02 * qr + v.G12 * qs + v.G22 * qt;
  double a = lambda * v.W * v.q;
  double v0, v1;
  pair(Gr, 1, v0, v1);
  a += Dm(0, i) * v0 + Dm(1, i) * v1;
  pair(Gs, 2, v0, v1);
  a += Dm(0, j) * v0 + Dm(1, j) * v1;
  pair(Gt, 4, v0, v1);
  a += Dm(0, k) * v0 + Dm(1, k) * v1;
  return a;
}

// block count of kernels B and C (and of their partials)
inline int coarse_row_blocks(const int Nrows) { return std::min((Nrows + RBS - 1) / RBS, RMAXB); }

// sum over the nb partials of one slot, in a fixed order, by the first wavefront
// (valid in lane 0 of wave 0)
// Partials of slots s0.. s0+NS-1 summed by one wavefront (threadIdx.x < 64):
// lane l adds entries l, l+64, l+128, ... in that order, then a wave sum. The
// loads of four steps of every slot are issued together (clamped indices;
// entries past nb add nothing) instead of one dependent load per step.
template <int NS>
__device__ inline void coarse_sum_slots(const ReduceWS* __restrict__ ws, const int s0, const int nb,
                                        double (&out)[NS]) {
  const int l = threadIdx.x;
#pragma unroll
  for (int k = 0; k < NS; ++k) out[k] = 0;
  for (int b0 = l; b0 < nb; b0 += 4 * 64) {
    double v[NS][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = min(b0 + 64 * u, nb - 1);
#pragma unroll
      for (int k = 0; k < NS; ++k) v[k][u] = ws->part[s0 + k][b];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (b0 + 64 * u < nb)
#pragma unroll
        for (int k = 0; k < NS; ++k) out[k] += v[k][u];
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) out[k] = wave_sum(out[k]);
}
__device__ inline double coarse_sum_slot(const ReduceWS* __restrict__ ws, const int slot, const int nb) {
  double a[1];
  coarse_sum_slots<1>(ws, slot, nb, a);
  return a[0];
}

__global__ void __launch_bounds__(256) coarse_ax_hex1_kernel(const int E, const int* __restrict__ G2L,
                                                             const double* __restrict__ wJ,
                                                             const double* __restrict__ ggeo,
                                                             const double* __restrict__ D, const double lambda,
                                                             const int first, const int nb,
                                                             const double* __restrict__ z,
                                                             const double* __restrict__ p, double* s,
                                                             const ReduceWS* __restrict__ ws,
                                                             double* __restrict__ AqL) {
  __shared__ double s_beta;
  __shared__ int s_stop;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = t >> 3, n = t & 7;
  const bool live = e < E;  // every lane takes part in the shuffles
  // loads that do not depend on beta go first
  Hex1Node v{};
  double zg = 0, pg = 0;
  if (live) {
    const int g = G2L[(size_t)e * 8 + n];
    if (g >= 0) {
      zg = z[g];
      if (!first) pg = p[g];
    }
    const double* G = ggeo + (size_t)e * 48 + n;
    v.G00 = G[0];
    v.G01 = G[8];
    v.G02 = G[16];
    v.G11 = G[24];
    v.G12 = G[32];
    v.G22 = G[40];
    v.W = wJ[(size_t)e * 8 + n];
  }
  if (threadIdx.x < 64) {
    double rr = 0, rz = 0;
    if (!first) {
      double g2[2];
      coarse_sum_slots<2>(ws, 1, nb, g2);
      rr = g2[0];
      rz = g2[1];
    }
    if (threadIdx.x == 0) {
      // STOP raised by an earlier launch: return without writing anything (one
      // thread reads the flag for the block, as in cg1_scalars)
      const bool done = s[8] != 0.0;
      int stop = done ? 1 : 0;
      double beta = 0.0;
      if (!first && !done) {
        if (s[7] > 0.0 && rr <= s[7])
          stop = 1;
        else
          beta = rz / s[0];  // linearSolverPCG.cpp:116-124
      }
      s_stop = stop;
      s_beta = beta;
      if (blockIdx.x == 0 && !done) {
        if (first) {
          s[3] = 0.0;
          s[1] = s[0];
        } else {
          s[5] = rr;
          if (stop) {
            s[8] = 1.0;
          } else {
            s[3] = beta;
            s[1] = rz;
          }
        }
      }
    }
  }
  __syncthreads();
  if (s_stop) return;  // the same decision in every block
  v.q = first ? zg : zg + s_beta * pg;
  const double a = hex1_apply(v, n, D, lambda);
  if (live) AqL[(size_t)e * 8 + n] = a;
}

__global__ void __launch_bounds__(RBS) coarse_gather_pap_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                                const int* __restrict__ colIds,
                                                                const double* __restrict__ AqL, const int first,
                                                                const double* __restrict__ z,
                                                                double* __restrict__ p, double* __restrict__ Ap,
                                                                double* s, ReduceWS* __restrict__ ws) {
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
      for (int k = 0; k < 8; ++k) c[k] = colIds[row_copy(a, len, k)];
#pragma unroll
      for (int k = 0; k < 8; ++k) xv[k] = AqL[c[k]];
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
  const double bs = block_sum(acc, s_w);
  if (threadIdx.x == 0) {
    ws->part[0][blockIdx.x] = bs;
    if (blockIdx.x == 0) s[0] = s[1];  // r.z of this iteration, for alpha
  }
}

__global__ void __launch_bounds__(RBS) coarse_update_kernel(const int N, const int nb, const double* __restrict__ invD,
                                                            const double* __restrict__ p,
                                                            const double* __restrict__ Ap, double* __restrict__ x,
                                                            double* __restrict__ r, double* __restrict__ z, double* s,
                                                            ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  __shared__ double s_alpha;
  if (s[8] != 0.0) return;
  const int stride = gridDim.x * blockDim.x;
  const int n0 = blockIdx.x * blockDim.x + threadIdx.x;
  // the first row's loads go before the p.Ap sum
  double x0 = 0, p0 = 0, r0 = 0, a0v = 0, d0 = 0;
  if (n0 < N) {
    x0 = x[n0];
    p0 = p[n0];
    r0 = r[n0];
    a0v = Ap[n0];
    d0 = invD[n0];
  }
  if (threadIdx.x < 64) {
    const double pap = coarse_sum_slot(ws, 0, nb);
    if (threadIdx.x == 0) {
      const double alpha = s[0] / pap;
      s_alpha = alpha;
      if (blockIdx.x == 0) {
        s[4] = pap;
        s[2] = alpha;
      }
    }
  }
  __syncthreads();
  const double alpha = s_alpha;
  double a0 = 0, a1 = 0;
  if (n0 < N) {
    x[n0] = x0 + alpha * p0;
    const double rn = r0 - alpha * a0v;
    r[n0] = rn;
    const double zn = 1.0 * d0 * rn;
    z[n0] = zn;
    a0 += rn * rn;
    a1 += rn * zn;
  }
  for (int n = n0 + stride; n < N; n += stride) {
    x[n] += alpha * p[n];
    const double rn = r[n] - alpha * Ap[n];
    r[n] = rn;
    const double zn = 1.0 * invD[n] * rn;
    z[n] = zn;
    a0 += rn * rn;
    a1 += rn * zn;
  }
  const double b0 = block_sum(a0, s_w);
  const double b1 = block_sum(a1, s_w);
  if (threadIdx.x == 0) {
    ws->part[1][blockIdx.x] = b0;
    ws->part[2][blockIdx.x] = b1;
    if (blockIdx.x == 0) s[9] += 1.0;
  }
}

// stopping test on the partials of the last queued iteration (one wavefront)
__global__ void __launch_bounds__(64) coarse_check_kernel(const int nb, double* s, const ReduceWS* __restrict__ ws) {
  if (s[8] != 0.0) return;
  const double rr = coarse_sum_slot(ws, 1, nb);
  if (threadIdx.x == 0) {
    s[5] = rr;
    if (s[7] > 0.0 && rr <= s[7]) s[8] = 1.0;
  }
}

// ---------------------------------------------------------------------------
// One global reduction per iteration: the Chronopoulos-Gear form of PCG
// (A. T. Chronopoulos and C. W. Gear, J. Comput. Appl. Math. 25 (1989)),
// which gives the PCG iterates in exact arithmetic. With u = invD r and
// w = A u:
//   gamma = r.u, delta = w.u (one fused reduction), beta = gamma/gamma_old,
//   alpha = gamma / (delta - beta gamma / alpha_old),
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = invD r.
// Every vector update is local to a node once alpha and beta are known, so it
// is folded into the next Ax kernel: each element lane recomputes its nodes'
// updated values from the old ones, the first local copy of each node (owner
// flag) writes them, and the Ax runs on the new u. p, s, r, u are double
// buffered by iteration parity (a block must not read a value another block
// has already replaced); x is read and written by its owner only. Two
// kernels per iteration:
//   cg1_ax:     partials -> gamma, delta, r.r; stop test; alpha, beta;
//               node updates; AqL = A_e u
//   cg1_gather: w = gather(AqL); partials of r.u, w.u, r.r
// Scalars: gamma and alpha of iteration i at [3P] and [3P+1] (P = i%2, so that
// the blocks reading the previous pair never see block 0 replace it), [2]
// beta, [5] r.r, [7] TOL, [8] STOP, [9] ITERS.
struct CG1Vec {
  double *p, *s, *r, *u;
};

// The step scalars of a single-reduction iteration from the previous launch's
// r.u, w.u, r.r partials: every block sums them in the same order and reaches
// the same alpha, beta and stop decision; block 0 records them in s.
__device__ inline void cg1_scalars(const int mode, const int P, const double tol2, const int nb, double* s,
                                   const ReduceWS* __restrict__ ws, double* s_ab, int& s_stop) {
  if (threadIdx.x < 64) {
    double ga = 0, de = 0, rr = 0;
    if (mode > 0) {
      double g3[3];
      coarse_sum_slots<3>(ws, 0, nb, g3);
      ga = g3[0];
      de = g3[1];
      rr = g3[2];
    }
    if (threadIdx.x == 0) {
      // STOP raised by an earlier launch: return without writing anything. One
      // thread reads the flag for the whole block (block 0 of the launch that
      // stops raises it while other blocks may still be starting), and the
      // decision reaches every wave through LDS after the barrier below.
      const bool done = s[8] != 0.0;
      int stop = done ? 1 : 0;
      double alpha = 0.0, beta = 0.0;
      if (done) {
      } else if (mode == 1) {
        // r0.r0 gives the relative threshold; a zero residual stops at once
        const double T = tol2 * rr;
        if (rr == 0.0) stop = 1;
        alpha = ga / de;
        if (blockIdx.x == 0) s[7] = (T > 0.0) ? T : 1e-300;
      } else if (mode == 2) {
        if (s[7] > 0.0 && rr <= s[7])
          stop = 1;
        else {
          const int q = 3 * (P ^ 1);  // the previous iteration's pair
          beta = ga / s[q];
          alpha = ga / (de - beta * ga / s[q + 1]);
        }
      }
      s_stop = stop;
      s_ab[0] = alpha;
      s_ab[1] = beta;
      if (blockIdx.x == 0 && mode > 0 && !done) {
        s[5] = rr;
        if (stop) {
          s[8] = 1.0;
        } else {
          s[3 * P] = ga;
          s[3 * P + 1] = alpha;
          s[2] = beta;
          s[9] += 1.0;
        }
      }
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(256) coarse_cg1_ax_kernel(
    const int E, const int* __restrict__ G2L, const unsigned char* __restrict__ owner, const double* __restrict__ wJ,
    const double* __restrict__ ggeo, const double* __restrict__ D, const double lambda, const int mode,
    const int P, const double tol2, const int nb, const double* __restrict__ invD, const double* __restrict__ w,
    double* __restrict__ x, const CG1Vec cur, const CG1Vec nxt, double* s, const ReduceWS* __restrict__ ws,
    double* __restrict__ AqL) {
  // mode 0: start (u = invD r, x = 0, no update); 1: first update (beta = 0); 2: later updates
  __shared__ double s_ab[2];
  __shared__ int s_stop;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = t >> 3, n = t & 7;
  const bool live = e < E;
  Hex1Node v{};
  int g = -1;
  bool own = false;
  double uo = 0, wo = 0, po = 0, so = 0, xo = 0, ro = 0, dg = 0;
  if (live) {
    g = G2L[(size_t)e * 8 + n];
    if (g >= 0) {
      own = owner[(size_t)e * 8 + n] != 0;
      ro = cur.r[g];
      dg = invD[g];
      if (mode > 0) {
        uo = cur.u[g];
        wo = w[g];
        if (own) xo = x[g];
        if (mode > 1) {
          po = cur.p[g];
          so = cur.s[g];
        }
      }
    }
    const double* G = ggeo + (size_t)e * 48 + n;
    v.G00 = G[0];
    v.G01 = G[8];
    v.G02 = G[16];
    v.G11 = G[24];
    v.G12 = G[32];
    v.G22 = G[40];
    v.W = wJ[(size_t)e * 8 + n];
  }
  cg1_scalars(mode, P, tol2, nb, s, ws, s_ab, s_stop);
  if (s_stop) return;  // the same decision in every block
  double un = 0.0;
  if (live && g >= 0) {
    if (mode == 0) {
      un = 1.0 * dg * ro;
      if (own) {
        x[g] = 0.0;
        cur.u[g] = un;  // mode 0 writes u0 in place: nothing reads u in this launch
      }
    } else {
      const double alpha = s_ab[0], beta = s_ab[1];
      const double pn = (mode == 1) ? uo : uo + beta * po;
      const double sn = (mode == 1) ? wo : wo + beta * so;
      const double rn = ro - alpha * sn;
      un = 1.0 * dg * rn;
      if (own) {
        x[g] = xo + alpha * pn;
        nxt.p[g] = pn;
        nxt.s[g] = sn;
        nxt.r[g] = rn;
        nxt.u[g] = un;
      }
    }
  }
  v.q = un;
  const double a = hex1_apply(v, n, D, lambda);
  if (live) AqL[(size_t)e * 8 + n] = a;
}

// w = gather(AqL); partials of r.u, w.u and r.r for the next cg1_ax
__global__ void __launch_bounds__(RBS) coarse_cg1_gather_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                                const int* __restrict__ colIds,
                                                                const double* __restrict__ AqL,
                                                                const double* __restrict__ r,
                                                                const double* __restrict__ u, double* __restrict__ w,
                                                                const double* s, ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  double ga = 0, de = 0, rr = 0;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < Nrows; g += gridDim.x * blockDim.x) {
    const int a = rowStarts[g], b = rowStarts[g + 1];
    double wg = 0;
    {
      const int len = b - a;
      int c[8];
      double xv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) c[k] = colIds[row_copy(a, len, k)];
#pragma unroll
      for (int k = 0; k < 8; ++k) xv[k] = AqL[c[k]];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < len) wg += xv[k];
      for (int j = a + 8; j < b; ++j) wg += AqL[colIds[j]];
    }
    w[g] = wg;
    const double rg = r[g], ug = u[g];
    ga += rg * ug;
    de += wg * ug;
    rr += rg * rg;
  }
  const double b0 = block_sum(ga, s_w);
  const double b1 = block_sum(de, s_w);
  const double b2 = block_sum(rr, s_w);
  if (threadIdx.x == 0) {
    ws->part[0][blockIdx.x] = b0;
    ws->part[1][blockIdx.x] = b1;
    ws->part[2][blockIdx.x] = b2;
  }
}

// stopping test on the r.r partials of the last cg1_gather (one wavefront);
// with mirror set, the scalars are then copied into pinned host memory mapped
// for the device, whether or not the solve had already stopped, so the host
// needs no separate copy after the chunk
__global__ void __launch_bounds__(64) coarse_cg1_check_kernel(const int nb, double* s, const ReduceWS* __restrict__ ws,
                                                              double* __restrict__ mirror) {
  if (s[8] == 0.0) {
    const double rr = coarse_sum_slot(ws, 2, nb);
    if (threadIdx.x == 0) {
      s[5] = rr;
      if (s[7] > 0.0 && rr <= s[7]) s[8] = 1.0;
    }
  }
  if (mirror && threadIdx.x == 0)
    for (int i = 0; i < LIBP_PCG_NSCALARS; ++i) mirror[i] = s[i];
}

// The single-reduction iteration on the assembled operator (CSR rows of the
// gathered dofs, columns ascending): one launch per iteration. A row's eight
// lanes recompute the updated u at every column from the previous iterate
// (read from the buffers of parity P) and form w = A u for the row; lane 0
// writes the row's updated vectors to parity P^1 and adds the row to the r.u,
// w.u, r.r partials. Every site evaluates the update with the same
// uncontracted expressions, so a column's u agrees bit for bit with the u its
// own row stores. Modes and scalars as in coarse_cg1_ax_kernel; mode 0 writes
// u and w of parity P in place (nothing reads them in that launch). The
// partials of the previous launch are read from wsIn and this launch's are
// written to wsOut, a different workspace: the blocks of one launch need not
// all be resident at once, so one block may finish before another starts.
constexpr int CSR_THREADS = 1024, CSR_LANES = 8, CSR_ROWS = CSR_THREADS / CSR_LANES, CSR_UNROLL = 32 / CSR_LANES;

__global__ void __launch_bounds__(CSR_THREADS) coarse_cg1_csr_kernel(
    const int Nrows, const int* __restrict__ rowStarts, const int* __restrict__ colIds,
    const double* __restrict__ vals, const int mode, const int P, const double tol2, const int nb,
    const double* __restrict__ invD, double* __restrict__ x, const CG1Vec cur, const CG1Vec nxt,
    double* __restrict__ wcur, double* __restrict__ wnxt, double* s, const ReduceWS* __restrict__ wsIn,
    ReduceWS* __restrict__ wsOut) {
#pragma clang fp contract(off)
  __shared__ double s_ab[2];
  __shared__ int s_stop;
  __shared__ double s_w[16];
  // u at gathered dof j after this iteration's update, from its loaded
  // r, w, s and invD
  auto unew = [&](double ro, double wo, double so, double dg, double alpha, double beta) -> double {
    if (mode == 0) return 1.0 * dg * ro;
    const double sn = (mode == 1) ? wo : wo + beta * so;
    const double rn = ro - alpha * sn;
    return 1.0 * dg * rn;
  };
  const int sub = threadIdx.x & (CSR_LANES - 1);
  double ga = 0, de = 0, rr = 0, alpha = 0, beta = 0;
  bool first = true;
  for (int g0 = blockIdx.x * CSR_ROWS; g0 < Nrows; g0 += gridDim.x * CSR_ROWS) {
    const int g = g0 + threadIdx.x / CSR_LANES;
    const bool row = g < Nrows;
    // every load of the pass is issued before the step scalars are formed:
    // none of them depends on the scalars
    int a = 0, b = 0;
    double uo = 0, wo = 0, ro = 0, dg = 0, po = 0, so = 0, xo = 0;
    if (row) {
      a = rowStarts[g];
      b = rowStarts[g + 1];
      if (sub == 0) {
        ro = cur.r[g];
        dg = invD[g];
        if (mode > 0) {
          uo = cur.u[g];
          wo = wcur[g];
          xo = x[g];
          if (mode > 1) {
            po = cur.p[g];
            so = cur.s[g];
          }
        }
      }
    }
    int c[CSR_UNROLL];
    double v[CSR_UNROLL], cr[CSR_UNROLL], cw[CSR_UNROLL], cs[CSR_UNROLL], cd[CSR_UNROLL];
    // the loads are unconditional, at positions clamped into the row (a row
    // holds at least its diagonal; a lane past the last row reads entry 0), so
    // they are all in flight together; entries past the row's end are dropped
    // from the sum through c[k] = -1
    int cj[CSR_UNROLL];
#pragma unroll
    for (int k = 0; k < CSR_UNROLL; ++k) {
      const int jk = a + sub + CSR_LANES * k;
      const int jc = min(jk, max(b - 1, 0));
      cj[k] = colIds[jc];
      const double vk = vals[jc];
      c[k] = (jk < b) ? cj[k] : -1;
      v[k] = (jk < b) ? vk : 0.0;
    }
#pragma unroll
    for (int k = 0; k < CSR_UNROLL; ++k) {
      const int j = cj[k];
      cr[k] = cur.r[j];
      cd[k] = invD[j];
      cw[k] = (mode > 0) ? wcur[j] : 0.0;
      cs[k] = (mode > 1) ? cur.s[j] : 0.0;
    }
    if (first) {
      cg1_scalars(mode, P, tol2, nb, s, wsIn, s_ab, s_stop);
      if (s_stop) return;  // the same decision in every block
      alpha = s_ab[0];
      beta = s_ab[1];
      first = false;
    }
    double acc = 0;
#pragma unroll
    for (int k = 0; k < CSR_UNROLL; ++k)
      if (c[k] >= 0) acc += v[k] * unew(cr[k], cw[k], cs[k], cd[k], alpha, beta);
    for (int jk = a + sub + CSR_UNROLL * CSR_LANES; jk < b; jk += CSR_LANES) {
      const int j = colIds[jk];
      acc += vals[jk] * unew(cur.r[j], mode > 0 ? wcur[j] : 0.0, mode > 1 ? cur.s[j] : 0.0, invD[j], alpha, beta);
    }
    // the row's sum over its lanes (xor butterfly: every lane holds the same value)
#pragma unroll
    for (int m = 1; m < CSR_LANES; m <<= 1) acc += __shfl_xor(acc, m, 64);
    if (row && sub == 0) {
      const double wv = acc;
      if (mode == 0) {
        const double un = unew(ro, 0.0, 0.0, dg, alpha, beta);
        x[g] = 0.0;
        cur.u[g] = un;
        wcur[g] = wv;
        ga += ro * un;
        de += wv * un;
        rr += ro * ro;
      } else {
        const double pn = (mode == 1) ? uo : uo + beta * po;
        const double sn = (mode == 1) ? wo : wo + beta * so;
        const double rn = ro - alpha * sn;
        const double un = 1.0 * dg * rn;
        x[g] = xo + alpha * pn;
        nxt.p[g] = pn;
        nxt.s[g] = sn;
        nxt.r[g] = rn;
        nxt.u[g] = un;
        wnxt[g] = wv;
        ga += rn * un;
        de += wv * un;
        rr += rn * rn;
      }
    }
  }
  const double b0 = block_sum(ga, s_w);
  const double b1 = block_sum(de, s_w);
  const double b2 = block_sum(rr, s_w);
  if (threadIdx.x == 0) {
    wsOut->part[0][blockIdx.x] = b0;
    wsOut->part[1][blockIdx.x] = b1;
    wsOut->part[2][blockIdx.x] = b2;
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
  // one row per thread: these vectors are small and the kernels latency-bound
  const int nb = coarse_row_blocks(Nrows);
  hipLaunchKernelGGL(coarse_ax_hex1_kernel, dim3((8 * (long)Nelements + 255) / 256), dim3(256), 0, st, Nelements,
                     GlobalToLocal, wJ, ggeo, D, lambda, first, nb, z, p, scalars, (const ReduceWS*)ws, AqL);
  hipLaunchKernelGGL(coarse_gather_pap_kernel, dim3(nb), dim3(RBS), 0, st, Nrows, rowStarts, colIds, AqL, first, z,
                     p, Ap, scalars, (ReduceWS*)ws);
  hipLaunchKernelGGL(coarse_update_kernel, dim3(nb), dim3(RBS), 0, st, Nrows, nb, invDiag, p, Ap, x, r, z, scalars,
                     (ReduceWS*)ws);
  return launch_status("pcg_jacobi_hex1_iteration");
}

extern "C" int libp_pcg_jacobi_hex1_check_f64(int Nrows, double* scalars, void* ws, void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!scalars || !ws) return arg_error("pcg_jacobi_hex1_check", "null pointer");
  hipLaunchKernelGGL(coarse_check_kernel, dim3(1), dim3(64), 0, as_stream(stream), coarse_row_blocks(Nrows), scalars,
                     (const ReduceWS*)ws);
  return launch_status("pcg_jacobi_hex1_check");
}

extern "C" int libp_pcg_jacobi_hex1_cg1_f64(int Nelements, const int* GlobalToLocal, const unsigned char* owner,
                                            const double* wJ, const double* ggeo, const double* D, double lambda,
                                            int Nrows, const int* rowStarts, const int* colIds, const double* invDiag,
                                            double tol, int mode, int parity, double* x, double* w, double* const* p2,
                                            double* const* s2, double* const* r2, double* const* u2, double* AqL,
                                            double* scalars, void* ws, void* stream) {
  if (Nelements <= 0 || Nrows <= 0) return LIBP_OK;
  if (!GlobalToLocal || !owner || !wJ || !ggeo || !D || !rowStarts || !colIds || !invDiag || !x || !w || !p2 ||
      !s2 || !r2 || !u2 || !AqL || !scalars || !ws)
    return arg_error("pcg_jacobi_hex1_cg1", "null pointer");
  if (mode < 0 || mode > 3 || (parity & ~1)) return arg_error("pcg_jacobi_hex1_cg1", "mode 0..3, parity 0 or 1");
  hipStream_t st = as_stream(stream);
  const int nb = coarse_row_blocks(Nrows);
  if (mode == 3) {
    hipLaunchKernelGGL(coarse_cg1_check_kernel, dim3(1), dim3(64), 0, st, nb, scalars, (const ReduceWS*)ws, nullptr);
    return launch_status("pcg_jacobi_hex1_cg1 (check)");
  }
  const int a = parity, b = parity ^ 1;
  const CG1Vec cur{p2[a], s2[a], r2[a], u2[a]}, nxt{p2[b], s2[b], r2[b], u2[b]};
  hipLaunchKernelGGL(coarse_cg1_ax_kernel, dim3((8 * (long)Nelements + 255) / 256), dim3(256), 0, st, Nelements,
                     GlobalToLocal, owner, wJ, ggeo, D, lambda, mode, parity, tol * tol, nb, invDiag, w, x, cur, nxt,
                     scalars, (const ReduceWS*)ws, AqL);
  // the gather reads the vectors this launch wrote: the start wrote u in place
  const CG1Vec& out = (mode == 0) ? cur : nxt;
  hipLaunchKernelGGL(coarse_cg1_gather_kernel, dim3(nb), dim3(RBS), 0, st, Nrows, rowStarts, colIds, AqL, out.r,
                     out.u, w, scalars, (ReduceWS*)ws);
  return launch_status("pcg_jacobi_hex1_cg1");
}

extern "C" int libp_pcg_jacobi_csr_cg1_f64(int Nrows, const int* rowStarts, const int* colIds, const double* vals,
                                           const double* invDiag, double tol, int mode, int parity, double* x,
                                           double* const* w2, double* const* p2, double* const* s2,
                                           double* const* r2, double* const* u2, double* scalars, void* ws,
                                           void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rowStarts || !colIds || !vals || !invDiag || !x || !w2 || !p2 || !s2 || !r2 || !u2 || !scalars || !ws)
    return arg_error("pcg_jacobi_csr_cg1", "null pointer");
  if (mode < 0 || mode > 3 || (parity & ~1)) return arg_error("pcg_jacobi_csr_cg1", "mode 0..3, parity 0 or 1");
  hipStream_t st = as_stream(stream);
  const int nb = (Nrows + CSR_ROWS - 1) / CSR_ROWS < RMAXB ? (Nrows + CSR_ROWS - 1) / CSR_ROWS : RMAXB;
  ReduceWS* w = (ReduceWS*)ws;  // two workspaces, used by parity
  const int a = parity, b = parity ^ 1;
  if (mode == 3) {
    // iteration (mode 1 or 2) of parity a wrote its partials to w[b]
    hipLaunchKernelGGL(coarse_cg1_check_kernel, dim3(1), dim3(64), 0, st, nb, scalars, (const ReduceWS*)(w + b),
                       nullptr);
    return launch_status("pcg_jacobi_csr_cg1 (check)");
  }
  const CG1Vec cur{p2[a], s2[a], r2[a], u2[a]}, nxt{p2[b], s2[b], r2[b], u2[b]};
  // mode 0 writes w[a], which iteration 0 (mode 1, parity a) reads
  hipLaunchKernelGGL(coarse_cg1_csr_kernel, dim3(nb), dim3(CSR_THREADS), 0, st, Nrows, rowStarts, colIds, vals, mode,
                     parity, tol * tol, nb, invDiag, x, cur, nxt, w2[a], w2[b], scalars, (const ReduceWS*)(w + a),
                     w + (mode == 0 ? a : b));
  return launch_status("pcg_jacobi_csr_cg1");
}

extern "C" int libp_pcg_jacobi_csr_cg1_check_f64(int Nrows, int parity, double* scalars, void* ws, double* hostMirror,
                                                 void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!scalars || !ws) return arg_error("pcg_jacobi_csr_cg1_check", "null pointer");
  if (parity & ~1) return arg_error("pcg_jacobi_csr_cg1_check", "parity 0 or 1");
  const int nb = (Nrows + CSR_ROWS - 1) / CSR_ROWS < RMAXB ? (Nrows + CSR_ROWS - 1) / CSR_ROWS : RMAXB;
  // iteration (mode 1 or 2) of parity p wrote its partials to workspace p ^ 1
  hipLaunchKernelGGL(coarse_cg1_check_kernel, dim3(1), dim3(64), 0, as_stream(stream), nb, scalars,
                     (const ReduceWS*)ws + (parity ^ 1), hostMirror);
  return launch_status("pcg_jacobi_csr_cg1_check");
}
