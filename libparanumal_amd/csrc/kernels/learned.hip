// Kernels of the learned per-element smoother around the MLP GEMMs (no
// reference counterpart): network features, assembly of the element
// corrections, the energy-optimal step that safeguards each application, and
// the loss / gradient kernels of training.
//
// Network input of element e (nIn = Np + 6*Nfp values): z = D^{-1} r at the
// element's own nodes (local order) then at the six neighbour traces (the
// layer of neighbour nodes next to each shared face, ordered by this element's
// face nodes), zero where there is no node (Dirichlet or no neighbour). The
// input is divided by its RMS s_e, and the output y_e is multiplied by s_e:
// the map r -> correction is positively homogeneous of degree one.
//
// Correction at gathered node g (rows of the masked gather):
//   c[g] = (1/|row g|) sum_{copies j} s_{e(j)} y[j] + omega D^{-1}_g r[g].
#include "gather_block.hpp"
#include "reduce.hpp"

namespace libp_amd {

typedef unsigned short u16;

__device__ inline u16 f2bf_l(float f) {
  unsigned u = __builtin_bit_cast(unsigned, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}

// Network input of NE elements, one wavefront: lane l holds the features
// f = l + 64 i (i < NPER) of each element as BF16 and returns the elements'
// RMS scales. e[n] < 0 stands for no element (zero input, scale 0). NP, NTR > 0
// fix Np and Ntr at compile time (the N=7 instance, 512 + 384): which map a
// lane's feature comes from is then known statically, and every index load of
// the lane is issued before the value loads that depend on them. NP = 0 takes
// the sizes at run time. The arithmetic of one element does not depend on NE.
template <int NPER, int NP, int NTR, int NE>
__device__ inline void element_inputs(const int (&e)[NE], const int l, const int Np_, const int Ntr_,
                                      const int* __restrict__ G2L, const int* __restrict__ traceIds,
                                      const float* __restrict__ r, const float* __restrict__ invD, u16 (&xb)[NE][NPER],
                                      float (&scale)[NE]) {
  const int Np = NP > 0 ? NP : Np_, Ntr = NP > 0 ? NTR : Ntr_;
  const int nIn = Np + Ntr;
  int gid[NE][NPER];
#pragma unroll
  for (int n = 0; n < NE; ++n)
#pragma unroll
    for (int i = 0; i < NPER; ++i) {
      const int f = l + 64 * i;
      int g = -1;
      if (e[n] >= 0) {
        if (f < Np) g = G2L[(long)e[n] * Np + f];
        else if (f < nIn) g = traceIds[(long)e[n] * Ntr + (f - Np)];
      }
      gid[n][i] = g;
    }
  float z[NE][NPER];
#pragma unroll
  for (int n = 0; n < NE; ++n)
#pragma unroll
    for (int i = 0; i < NPER; ++i) {
      const int g = gid[n][i];
      const int gc = max(g, 0);  // unconditional loads: a missing node reads entry 0 and takes 0
      const float dv = invD[gc], rv = r[gc];
      z[n][i] = (g >= 0) ? dv * rv : 0.f;
    }
#pragma unroll
  for (int n = 0; n < NE; ++n) {
    double ss = 0;
#pragma unroll
    for (int i = 0; i < NPER; ++i) ss += (double)z[n][i] * z[n][i];
    ss = wave_sum(ss);
    const float s = (float)sqrt(ss / nIn);
    const float inv = (s > 0.f) ? 1.f / s : 0.f;
#pragma unroll
    for (int i = 0; i < NPER; ++i) xb[n][i] = f2bf_l(z[n][i] * inv);
    scale[n] = s;
  }
}

// One wavefront per element, four per block.
template <int NPER, int NP, int NTR>
__global__ void __launch_bounds__(256) learned_features_kernel(const int E, const int Np_, const int Ntr_,
                                                               const int* __restrict__ G2L,
                                                               const int* __restrict__ traceIds,
                                                               const float* __restrict__ r,
                                                               const float* __restrict__ invD, u16* __restrict__ X,
                                                               const long ldx, float* __restrict__ scale) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (e >= E) return;
  const int nIn = NP > 0 ? NP + NTR : Np_ + Ntr_;
  const int ee[1] = {e};
  u16 xb[1][NPER];
  float s[1];
  element_inputs<NPER, NP, NTR, 1>(ee, l, Np_, Ntr_, G2L, traceIds, r, invD, xb, s);
#pragma unroll
  for (int i = 0; i < NPER; ++i) {
    const int f = l + 64 * i;
    if (f < nIn) X[(long)e * ldx + f] = xb[0][i];
  }
  if (l == 0) scale[e] = s[0];
}

// Linear layer of the N=7 instance fused with its input: Y[e][n] =
// s_e * sum_k X[e][k] W[n][k] for 896 inputs and 512 outputs, where X is the
// BF16 input of element_inputs and s_e its scale. The products, their grouping
// into v_mfma_f32_32x32x16_bf16 steps over k = 16t..16t+15 in ascending t, the
// operand lanes and the epilogue are those of gemm_bf16_kernel with
// libp_mlp_linear_fwd_scaled, so Y is the same to the bit; X never goes to
// memory.
//
// One workgroup of eight wavefronts per 64 elements. Phase 1: each wavefront
// builds the inputs of eight elements, two at a time, into LDS rows of 904
// BF16 (the 8-element pad puts consecutive rows four banks apart, so the
// MFMA operand reads are conflict-free). Phase 2: wavefront w forms columns
// 64w..64w+63 of all 64 rows (2x2 MFMA blocks of 32x32). Its weights come from
// the packed copy (learned_pack_weights_kernel: one 1 KiB contiguous block
// per 32 columns and 16 inputs, in the MFMA's lane order) through a ring of
// LL_PF steps of register prefetch; the 917 KB of packed weights stay in L2.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int LL_ROWS = 64, LL_NP = 512, LL_NTR = 384, LL_K = LL_NP + LL_NTR, LL_N = 512;
constexpr int LL_LDA = LL_K + 8, LL_STEPS = LL_K / 16, LL_PF = 8;
static_assert(LL_STEPS % LL_PF == 0, "prefetch ring must divide the steps");

__global__ void __launch_bounds__(512, 1) learned_linear_layer_kernel(const int E, const int* __restrict__ G2L,
                                                                       const int* __restrict__ traceIds,
                                                                       const float* __restrict__ r,
                                                                       const float* __restrict__ invD,
                                                                       const u32x4* __restrict__ Wp,
                                                                       float* __restrict__ Y, const long ldy,
                                                                       float* __restrict__ scale) {
  __shared__ __attribute__((aligned(16))) u16 sA[LL_ROWS * LL_LDA];
  __shared__ float sS[LL_ROWS];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const long e0 = (long)blockIdx.x * LL_ROWS;

  // phase 1: rows 8w..8w+7
#pragma unroll 1
  for (int q = 0; q < 8; q += 2) {
    const int row0 = 8 * w + q;
    const int ee[2] = {e0 + row0 < E ? (int)(e0 + row0) : -1, e0 + row0 + 1 < E ? (int)(e0 + row0 + 1) : -1};
    u16 xb[2][14];
    float sc[2];
    element_inputs<14, LL_NP, LL_NTR, 2>(ee, l, LL_NP, LL_NTR, G2L, traceIds, r, invD, xb, sc);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
#pragma unroll
      for (int i = 0; i < 14; ++i) sA[(row0 + n) * LL_LDA + l + 64 * i] = xb[n][i];
      if (l == 0) {
        sS[row0 + n] = sc[n];
        if (scale && ee[n] >= 0) scale[ee[n]] = sc[n];
      }
    }
  }
  __syncthreads();

  // phase 2: columns 64w..64w+63
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 16; ++k) acc[i][j][k] = 0.f;
  const int lr = l & 31, lk = 8 * (l >> 5);
  const u32x4* wp0 = Wp + (long)(2 * w) * LL_STEPS * 64 + l;
  const u32x4* wp1 = wp0 + LL_STEPS * 64;
  const u16* a0p = sA + lr * LL_LDA + lk;
  const u16* a1p = a0p + 32 * LL_LDA;
  u32x4 bq[LL_PF][2];
#pragma unroll
  for (int p = 0; p < LL_PF; ++p) {
    bq[p][0] = wp0[p * 64];
    bq[p][1] = wp1[p * 64];
  }
#pragma unroll 1
  for (int s0 = 0; s0 < LL_STEPS; s0 += LL_PF) {
#pragma unroll
    for (int p = 0; p < LL_PF; ++p) {
      const int st = s0 + p;
      const bf16x8 b0 = __builtin_bit_cast(bf16x8, bq[p][0]);
      const bf16x8 b1 = __builtin_bit_cast(bf16x8, bq[p][1]);
      if (st + LL_PF < LL_STEPS) {
        bq[p][0] = wp0[(st + LL_PF) * 64];
        bq[p][1] = wp1[(st + LL_PF) * 64];
      }
      const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(a0p + st * 16);
      const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(a1p + st * 16);
      acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[1][1], 0, 0, 0);
    }
  }

  // epilogue: s_e * (sum + 0), the bias-free form of the GEMM's scaled epilogue
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = 64 * w + 32 * j + lr;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int row = 32 * i + 8 * (k >> 2) + 4 * (l >> 5) + (k & 3);
        if (e0 + row >= E) continue;
        const float o = acc[i][j][k] + 0.f;
        Y[(e0 + row) * ldy + col] = sS[row] * o;
      }
    }
}

// Packed weights of the fused layer: 16-byte chunk (cb*K/16 + t)*64 + lane
// holds W[32 cb + lane%32][16 t + 8 (lane/32) .. +7], the B operand that lane
// supplies to step t of the MFMA for the 32 columns of block cb.
__global__ void __launch_bounds__(256) learned_pack_weights_kernel(const int nChunks, const int ksteps,
                                                                   const u16* __restrict__ W, const long ldw,
                                                                   u32x4* __restrict__ Wp) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nChunks) return;
  const int lane = c & 63, st = (c >> 6) % ksteps, cb = (c >> 6) / ksteps;
  const u16* src = W + (long)(32 * cb + (lane & 31)) * ldw + 16 * st + 8 * (lane >> 5);
  u32x4 v;
#pragma unroll
  for (int h = 0; h < 4; ++h) v[h] = (unsigned)src[2 * h] | ((unsigned)src[2 * h + 1] << 16);
  Wp[c] = v;
}

// c[g] = invLen * sum_j s_e Y[j] + omega * invD[g] r[g]; omega read from device
// memory, the term skipped when omega is null. SHIFT = log2(Np) when Np is a
// power of two (the element of a local id is then a shift), else -1.
// ACC: c += the correction (the smoother's x += c in the same pass; the sum
// rounds once, as the separate axpy it replaces does). SCALED = false: Y
// already holds s_e y (libp_mlp_linear_fwd_scaled forms the same products in
// its epilogue), so each copy is one load.
template <int SHIFT, bool ACC = false, bool SCALED = true>
__global__ void __launch_bounds__(GB_THREADS) learned_assemble_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                                      const int* __restrict__ colIds, const int Np,
                                                                      const float* __restrict__ Y, const long ldy,
                                                                      const float* __restrict__ scale,
                                                                      const float* __restrict__ omega,
                                                                      const float* __restrict__ r,
                                                                      const float* __restrict__ invD,
                                                                      float* __restrict__ c) {
  // rows in row blocks (gather_block_sums_by): the copies' s_e * y values staged in
  // LDS, each row summed from zero in ascending order
  __shared__ float s_v[GB_CAP];
  const float om = omega ? *omega : 0.f;
  auto copy_value = [=](int id) {
    const int e = SHIFT >= 0 ? (id >> SHIFT) : id / Np;
    const int n = id - e * Np;
    return SCALED ? scale[e] * Y[(long)e * ldy + n] : Y[(long)e * ldy + n];
  };
  const long nrb = gather_row_blocks(Nrows);
  for (long rb = blockIdx.x; rb < nrb; rb += gridDim.x) {
    GatherRows<float> g;
    gather_block_sums_by(rb, Nrows, rowStarts, colIds, copy_value, s_v, g);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (!g.in[u]) continue;
      const long row = g.row[u];
      const int len = g.len[u];
      float v = (len > 0 ? g.val[u] / (float)len : 0.f);
      if (omega) v += om * invD[row] * r[row];
      if (ACC)
        c[row] = 1.f * v + 1.f * c[row];
      else
        c[row] = v;
    }
  }
}

/* Energy-optimal step along the correction c for the residual res:
 *   alpha = (c . res) / (c . A c).
 * sc[0] = c.res, sc[1] = c.Ac, sc[2] = alpha, sc[3] += 1 when the step is
 * rejected (non-finite or c.Ac <= 0: alpha = 0 then). */
__global__ void __launch_bounds__(RBS) learned_alpha_kernel(const int N, const float* __restrict__ c,
                                                            const float* __restrict__ res,
                                                            const float* __restrict__ Ac, double* __restrict__ sc,
                                                            ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double a0 = 0, a1 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const double cn = c[n];
    a0 += cn * res[n];
    a1 += cn * Ac[n];
  }
  double v[2] = {a0, a1};
  store_block_partials<2>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) learned_alpha_finish_kernel(const int nb, double* __restrict__ sc,
                                                                   const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double out[2];
  sum_block_partials<2>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    sc[0] = out[0];
    sc[1] = out[1];
    const double alpha = out[0] / out[1];
    const bool ok = (out[1] > 0.0) && __builtin_isfinite(alpha) && __builtin_isfinite(out[0]);
    sc[2] = ok ? alpha : 0.0;
    if (!ok) sc[3] += 1.0;
  }
}

// x = (xIsZero ? 0 : x) + alpha c ; res -= alpha Ac (when res != null)
__global__ void __launch_bounds__(256) learned_update_kernel(const int N, const int xIsZero,
                                                             const double* __restrict__ sc,
                                                             const float* __restrict__ c,
                                                             const float* __restrict__ Ac, float* __restrict__ x,
                                                             float* __restrict__ res) {
  const float alpha = (float)sc[2];
  // a rejected step (alpha = 0) leaves x and res untouched even where c is not finite
  const bool take = alpha != 0.f;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const float xn = xIsZero ? 0.f : x[n];
    x[n] = take ? xn + alpha * c[n] : xn;
    if (res && take) res[n] -= alpha * Ac[n];
  }
}

/* Training loss of one sample and its gradient with respect to c, with
 * d = eh - c and Ad = A d (A symmetric):
 *   J = d.Ad / nrm,  dJ/dc = -2 Ad / nrm,  dJ/domega = sum_g dJ/dc[g] invD[g] r[g].
 * Writes gc, stat[0] = J, stat[1] += J, grad[ofsOmega] = dJ/domega. */
__global__ void __launch_bounds__(RBS) learned_loss_kernel(const int N, const float* __restrict__ d,
                                                           const float* __restrict__ Ad,
                                                           const float* __restrict__ r,
                                                           const float* __restrict__ invD,
                                                           const double* __restrict__ nrm, float* __restrict__ gc,
                                                           double* __restrict__ stat, float* __restrict__ gradOmega,
                                                           ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  const double scaleG = -2.0 / *nrm;
  double a0 = 0, a1 = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const double adn = Ad[n];
    const double g = scaleG * adn;
    gc[n] = (float)g;
    a0 += (double)d[n] * adn;
    a1 += g * (double)invD[n] * (double)r[n];
  }
  double v[2] = {a0, a1};
  store_block_partials<2>(ws, v, s_w);
}

__global__ void __launch_bounds__(RBS) learned_loss_finish_kernel(const int nb, const double* __restrict__ nrm,
                                                                  double* __restrict__ stat,
                                                                  float* __restrict__ gradOmega,
                                                                  const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  double out[2];
  sum_block_partials<2>(ws, nb, out, s_w);
  if (threadIdx.x == 0) {
    const double J = out[0] / *nrm;
    stat[0] = J;
    stat[1] += J;
    if (gradOmega) *gradOmega = (float)out[1];
  }
}

// dY[e][n] = bf16(gc[g] * s_e * invLen[g]) at g = G2L[e*Np+n] >= 0, else 0
__global__ void __launch_bounds__(256) learned_scatter_grad_kernel(const int E, const int Np,
                                                                   const int* __restrict__ G2L,
                                                                   const float* __restrict__ gc,
                                                                   const float* __restrict__ invLen,
                                                                   const float* __restrict__ scale,
                                                                   u16* __restrict__ dY, const long lddy) {
  const long NL = (long)E * Np;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < NL; i += (long)gridDim.x * blockDim.x) {
    const int e = (int)(i / Np), n = (int)(i - (long)e * Np);
    const int g = G2L[i];
    const int gs = max(g, 0);
    const float gv = gc[gs], lv = invLen[gs];
    const float v = (g >= 0) ? gv * scale[e] * lv : 0.f;
    dY[(long)e * lddy + n] = f2bf_l(v);
  }
}

// Least-squares target of the linear smoother: T[e][n] = eh[g] / s_e at the
// element's nodes (zero at masked nodes), the element-local error that the
// output y_e = W x_e should reproduce before it is multiplied by s_e.
__global__ void __launch_bounds__(256) learned_element_target_kernel(const int E, const int Np,
                                                                     const int* __restrict__ G2L,
                                                                     const float* __restrict__ eh,
                                                                     const float* __restrict__ scale,
                                                                     u16* __restrict__ T, const long ldt) {
  const long NL = (long)E * Np;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < NL; i += (long)gridDim.x * blockDim.x) {
    const int e = (int)(i / Np), n = (int)(i - (long)e * Np);
    const int g = G2L[i];
    const float s = scale[e];
    const float v = (g >= 0 && s > 0.f) ? eh[g] / s : 0.f;
    T[(long)e * ldt + n] = f2bf_l(v);
  }
}

// x[i] = uniform(-1, 1) from splitmix64 of (seed, i): white-noise error
// samples for the linear smoother's training (a pure function of seed and i)
__global__ void __launch_bounds__(256) random_uniform_kernel(const long N, const unsigned long long seed,
                                                             float* __restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    unsigned long long z = seed + 0x9e3779b97f4a7c15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    x[i] = (float)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
  }
}

// invLen[g] = 1 / (rowStarts[g+1] - rowStarts[g])
__global__ void __launch_bounds__(256) inv_row_length_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                             float* __restrict__ invLen) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < Nrows; g += gridDim.x * blockDim.x) {
    const int len = rowStarts[g + 1] - rowStarts[g];
    invLen[g] = len > 0 ? 1.f / (float)len : 0.f;
  }
}

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_learned_features_f32(int Nelements, int Np, int Ntr, const int* GlobalToLocal,
                                         const int* traceIds, const float* r, const float* invDiag,
                                         unsigned short* X, long ldx, float* scale, void* stream) {
  if (Nelements <= 0) return LIBP_OK;
  if (!GlobalToLocal || (Ntr > 0 && !traceIds) || !r || !invDiag || !X || !scale)
    return arg_error("learned_features", "null pointer");
  const int nIn = Np + Ntr;
  if (Np <= 0 || Ntr < 0 || ldx < nIn) return arg_error("learned_features", "bad size");
  const int nper = (nIn + 63) / 64;
  dim3 grid((Nelements + 3) / 4), block(256);
  hipStream_t s = as_stream(stream);
  if (Np == 512 && Ntr == 384) {
    hipLaunchKernelGGL((learned_features_kernel<14, 512, 384>), grid, block, 0, s, Nelements, Np, Ntr, GlobalToLocal,
                       traceIds, r, invDiag, X, ldx, scale);
    return launch_status("learned_features");
  }
#define LF(NPER)                                                                                                  \
  case NPER:                                                                                                      \
    hipLaunchKernelGGL((learned_features_kernel<NPER, 0, 0>), grid, block, 0, s, Nelements, Np, Ntr,             \
                       GlobalToLocal, traceIds, r, invDiag, X, ldx, scale);                                       \
    break;
  switch (nper) {
    LF(1) LF(2) LF(3) LF(4) LF(5) LF(6) LF(7) LF(8) LF(9) LF(10) LF(11) LF(12) LF(13) LF(14) LF(15) LF(16)
    default:
      return arg_error("learned_features", "more than 1024 features per element");
  }
#undef LF
  return launch_status("learned_features");
}

extern "C" int libp_learned_pack_weights_bf16(int Nout, int K, const unsigned short* W, long ldw,
                                              unsigned short* Wpacked, void* stream) {
  if (Nout <= 0 || K <= 0) return LIBP_OK;
  if (!W || !Wpacked) return arg_error("learned_pack_weights", "null pointer");
  if (Nout % 32 || K % 16 || ldw < K) return arg_error("learned_pack_weights", "Nout must be a multiple of 32, K of 16");
  if (((uintptr_t)Wpacked & 15) != 0) return arg_error("learned_pack_weights", "packed weights must be 16-byte aligned");
  const int nChunks = (Nout / 32) * (K / 16) * 64;
  hipLaunchKernelGGL(learned_pack_weights_kernel, dim3((nChunks + 255) / 256), dim3(256), 0, as_stream(stream),
                     nChunks, K / 16, W, ldw, reinterpret_cast<u32x4*>(Wpacked));
  return launch_status("learned_pack_weights");
}

extern "C" int libp_learned_linear_layer_f32(int Nelements, int Np, int Ntr, int Nout, const int* GlobalToLocal,
                                             const int* traceIds, const float* r, const float* invDiag,
                                             const unsigned short* Wpacked, float* Y, long ldy, float* scale,
                                             void* stream) {
  if (Nelements <= 0) return LIBP_OK;
  if (Np != LL_NP || Ntr != LL_NTR || Nout != LL_N)
    return arg_error("learned_linear_layer", "only the N=7 shape (512 + 384 inputs, 512 outputs)");
  if (!GlobalToLocal || !traceIds || !r || !invDiag || !Wpacked || !Y)
    return arg_error("learned_linear_layer", "null pointer");
  if (ldy < Nout) return arg_error("learned_linear_layer", "bad size");
  if (((uintptr_t)Wpacked & 15) != 0) return arg_error("learned_linear_layer", "packed weights must be 16-byte aligned");
  hipLaunchKernelGGL(learned_linear_layer_kernel, dim3((Nelements + LL_ROWS - 1) / LL_ROWS), dim3(512), 0,
                     as_stream(stream), Nelements, GlobalToLocal, traceIds, r, invDiag,
                     reinterpret_cast<const u32x4*>(Wpacked), Y, ldy, scale);
  return launch_status("learned_linear_layer");
}

template <bool ACC, bool SCALED>
static void assemble_launch(dim3 grid, dim3 block, hipStream_t st, int Nrows, const int* rowStarts,
                            const int* colIds, int Np, const float* Y, long ldy, const float* scale,
                            const float* omega, const float* r, const float* invDiag, float* c) {
  if (Np == 512)
    hipLaunchKernelGGL((learned_assemble_kernel<9, ACC, SCALED>), grid, block, 0, st, Nrows, rowStarts, colIds, Np,
                       Y, ldy, scale, omega, r, invDiag, c);
  else if (Np == 64)
    hipLaunchKernelGGL((learned_assemble_kernel<6, ACC, SCALED>), grid, block, 0, st, Nrows, rowStarts, colIds, Np,
                       Y, ldy, scale, omega, r, invDiag, c);
  else
    hipLaunchKernelGGL((learned_assemble_kernel<-1, ACC, SCALED>), grid, block, 0, st, Nrows, rowStarts, colIds, Np,
                       Y, ldy, scale, omega, r, invDiag, c);
}

// scale null: Y holds the scaled outputs already
template <bool ACC>
static int assemble(int Nrows, const int* rowStarts, const int* colIds, int Np, const float* Y, long ldy,
                    const float* scale, const float* omega, const float* r, const float* invDiag, float* c,
                    void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rowStarts || !colIds || !Y || !r || !invDiag || !c) return arg_error("learned_assemble", "null pointer");
  if (Np <= 0 || ldy < Np) return arg_error("learned_assemble", "bad size");
  const dim3 grid(row_block_grid(Nrows)), block(GB_THREADS);
  hipStream_t st = as_stream(stream);
  if (scale)
    assemble_launch<ACC, true>(grid, block, st, Nrows, rowStarts, colIds, Np, Y, ldy, scale, omega, r, invDiag, c);
  else
    assemble_launch<ACC, false>(grid, block, st, Nrows, rowStarts, colIds, Np, Y, ldy, scale, omega, r, invDiag, c);
  return launch_status("learned_assemble");
}

extern "C" int libp_learned_assemble_f32(int Nrows, const int* rowStarts, const int* colIds, int Np, const float* Y,
                                         long ldy, const float* scale, const float* omega, const float* r,
                                         const float* invDiag, float* c, void* stream) {
  return assemble<false>(Nrows, rowStarts, colIds, Np, Y, ldy, scale, omega, r, invDiag, c, stream);
}
extern "C" int libp_learned_assemble_add_f32(int Nrows, const int* rowStarts, const int* colIds, int Np,
                                             const float* Y, long ldy, const float* scale, const float* omega,
                                             const float* r, const float* invDiag, float* c, void* stream) {
  return assemble<true>(Nrows, rowStarts, colIds, Np, Y, ldy, scale, omega, r, invDiag, c, stream);
}

extern "C" int libp_learned_alpha_f32(int N, const float* c, const float* res, const float* Ac, double* sc, void* ws,
                                      void* stream) {
  if (!c || !res || !Ac || !sc || !ws) return arg_error("learned_alpha", "null pointer");
  hipLaunchKernelGGL(learned_alpha_kernel, dim3(reduce_blocks(N)), dim3(RBS), 0, as_stream(stream), N, c, res, Ac, sc,
                     (ReduceWS*)ws);
  hipLaunchKernelGGL(learned_alpha_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)reduce_blocks(N), sc,
                     (const ReduceWS*)ws);
  return launch_status("learned_alpha");
}

extern "C" int libp_learned_update_f32(int N, int xIsZero, const double* sc, const float* c, const float* Ac, float* x,
                                       float* res, void* stream) {
  if (N <= 0) return LIBP_OK;
  if (!sc || !c || !x || (res && !Ac)) return arg_error("learned_update", "null pointer");
  hipLaunchKernelGGL(learned_update_kernel, dim3(stream_blocks(N, 256)), dim3(256), 0, as_stream(stream), N, xIsZero, sc,
                     c, Ac, x, res);
  return launch_status("learned_update");
}

extern "C" int libp_learned_loss_f32(int N, const float* d, const float* Ad, const float* r, const float* invDiag,
                                     const double* nrm, float* gc, double* stat, float* gradOmega, void* ws,
                                     void* stream) {
  if (!d || !Ad || !r || !invDiag || !nrm || !gc || !stat || !ws) return arg_error("learned_loss", "null pointer");
  hipLaunchKernelGGL(learned_loss_kernel, dim3(reduce_blocks(N)), dim3(RBS), 0, as_stream(stream), N, d, Ad, r, invDiag,
                     nrm, gc, stat, gradOmega, (ReduceWS*)ws);
  hipLaunchKernelGGL(learned_loss_finish_kernel, dim3(1), dim3(RBS), 0, as_stream(stream), (int)reduce_blocks(N), nrm,
                     stat, gradOmega, (const ReduceWS*)ws);
  return launch_status("learned_loss");
}

extern "C" int libp_learned_scatter_grad_f32(int Nelements, int Np, const int* GlobalToLocal, const float* gc,
                                             const float* invLen, const float* scale, unsigned short* dY, long lddy,
                                             void* stream) {
  if (Nelements <= 0) return LIBP_OK;
  if (!GlobalToLocal || !gc || !invLen || !scale || !dY) return arg_error("learned_scatter_grad", "null pointer");
  if (lddy < Np) return arg_error("learned_scatter_grad", "bad size");
  hipLaunchKernelGGL(learned_scatter_grad_kernel, dim3(stream_blocks((long)Nelements * Np, 256)), dim3(256), 0,
                     as_stream(stream), Nelements, Np, GlobalToLocal, gc, invLen, scale, dY, lddy);
  return launch_status("learned_scatter_grad");
}

extern "C" int libp_inv_row_length_f32(int Nrows, const int* rowStarts, float* invLen, void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rowStarts || !invLen) return arg_error("inv_row_length", "null pointer");
  hipLaunchKernelGGL(inv_row_length_kernel, dim3(stream_blocks(Nrows, 256)), dim3(256), 0, as_stream(stream), Nrows,
                     rowStarts, invLen);
  return launch_status("inv_row_length");
}

extern "C" int libp_learned_element_target_f32(int Nelements, int Np, const int* GlobalToLocal, const float* eh,
                                               const float* scale, unsigned short* T, long ldt, void* stream) {
  if (Nelements <= 0) return LIBP_OK;
  if (!GlobalToLocal || !eh || !scale || !T) return arg_error("learned_element_target", "null pointer");
  if (Np <= 0 || ldt < Np) return arg_error("learned_element_target", "bad size");
  hipLaunchKernelGGL(learned_element_target_kernel, dim3(stream_blocks((long)Nelements * Np, 256)), dim3(256), 0,
                     as_stream(stream), Nelements, Np, GlobalToLocal, eh, scale, T, ldt);
  return launch_status("learned_element_target");
}

extern "C" int libp_random_uniform_f32(long N, unsigned long long seed, float* x, void* stream) {
  if (N <= 0) return LIBP_OK;
  if (!x) return arg_error("random_uniform", "null pointer");
  hipLaunchKernelGGL(random_uniform_kernel, dim3(stream_blocks(N, 256)), dim3(256), 0, as_stream(stream), N, seed, x);
  return launch_status("random_uniform");
}
