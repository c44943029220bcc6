// N=15 hex Ax with the 16x16 tensor contractions on FP64 MFMA
// (v_mfma_f64_16x16x4f64, wave64).
//
// Operand layout of v_mfma_f64_16x16x4f64 (C += A[16x4] B[4x16]), verified on
// gfx950: lane l supplies a = A[l%16][l/16] and b = B[l/16][l%16], and holds
// c[r] = C[4r + l/16][l%16]. A 16x16x16 product is four MFMAs over K chunks
// 4c..4c+3.
//
// One element (16^3 nodes) per workgroup of four wavefronts. With li = l%16,
// lg = l/16, a lane of the k-slice phase owns the nodes (i=li, j=4r+lg, k),
// r=0..3 ("L2" layout), which is both the MFMA output layout and the B
// operand layout for contractions over j. Contractions over i need the A
// operand in the transposed ("L1") layout, obtained through LDS.
//
//   phase T : waves take j-slices, qt = D q along k          (MFMA, LDS in/out)
//   phase K : waves take k-slices: qr = Q D^T, qs = D Q       (MFMA)
//             G factors (streamed, non-temporal), Gqs D^T-part, Gqr via LDS
//             transpose, partial Aq = r-part + s-part + lambda wJ q -> LDS
//   phase T2: waves take j-slices, add D^T Gqt along k, write Aq
//
// LDS: q / partial Aq (32 KiB), qt / Gqt (32 KiB), per-wave transpose tile.
// Reference math: okl/ellipticAxHex3D.okl:156-295 (same operator; the
// summation order differs, results agree to rounding).
//
// PAP (the PCG form): the workgroup also stores sum over its element's nodes
// of q_L * (A_L q)_L at pap[blockIdx.x], its share of p.Ap = p^T Q^T A_L Q p
// for q = p, and returns at once while scal[8] (STOP) is set. The sum is
// formed in phase K from values in registers: with A_L q = P + D_t^T Gqt (P
// the r, s and lambda parts, D_t the derivative along k),
// q . A_L q = q . P + (D_t q) . Gqt = sum of q P + qt Gqt over the nodes, so
// no q has to outlive phase K. The E-vector is the same, bit for bit.
#include "reduce.hpp"

namespace libp_amd {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <bool PAP>
__global__ void __launch_bounds__(256, 1) ax_hex_mfma_n15_kernel(const int* __restrict__ elementList,
                                                                 const int* __restrict__ G2L,
                                                                 const double* __restrict__ wJ,
                                                                 const double* __restrict__ ggeo,
                                                                 const double* __restrict__ D, const double lambda,
                                                                 const double* __restrict__ q,
                                                                 double* __restrict__ Aq,
                                                                 double* __restrict__ pap = nullptr,
                                                                 const double* __restrict__ scal = nullptr) {
  // a*b + c contracts only within one source expression, so both
  // instantiations round the E-vector alike
#pragma clang fp contract(on)
  constexpr int Nq = 16, NqNq = 256, Np = 4096;
  if (PAP && scal[8] != 0.0) return;  // a PCG step after convergence (STOP)
  __shared__ double s_q[Np];    // q, then r+s+lambda partial of Aq
  __shared__ double s_t[Np];    // qt, then Gqt
  __shared__ double s_x[4][NqNq];  // per-wave transpose tile
  const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lg = l >> 4;
  const int element = elementList ? elementList[blockIdx.x] : blockIdx.x;
  const size_t eb = (size_t)element * Np;

  // D fragments: dA[c] = D[li][4c+lg] (A operand of D, B operand of D^T),
  //              dT[c] = D[4c+lg][li] (A operand of D^T, B operand of D)
  double dA[4], dT[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    dA[c] = D[li * Nq + 4 * c + lg];
    dT[c] = D[(4 * c + lg) * Nq + li];
  }

  // gather q into LDS: all 16 index loads in flight, then all 16 value loads
  {
    int id[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) id[u] = __builtin_nontemporal_load(G2L + eb + t + 256 * u);
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const double qv = q[max(id[u], 0)];  // unconditional: masked nodes read q[0] and take 0
      v[u] = (id[u] >= 0) ? qv : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) s_q[t + 256 * u] = v[u];
  }
  __syncthreads();

  // phase K's geometric factors for its first k-slice are loaded here, so
  // that their HBM latency overlaps phase T; each later slice's are loaded
  // one slice ahead
  const double* gbase = ggeo + (size_t)element * 6 * Np;
  const double* wbase = wJ + eb;
  double Gn[6][4], Wn[4];
  auto loadG = [&](int k) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = k * NqNq + (4 * r + lg) * Nq + li;
#pragma unroll
      for (int g = 0; g < 6; ++g) Gn[g][r] = __builtin_nontemporal_load(gbase + g * Np + n);
      Wn[r] = __builtin_nontemporal_load(wbase + n);
    }
  };
  loadG(4 * w);

  // phase T: qt(i,j,k) = sum_m D[k][m] q(i,j,m) for j = 4w..4w+3
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int j = 4 * w + jj;
    d4 acc = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = mfma16(dA[c], s_q[(4 * c + lg) * NqNq + j * Nq + li], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) s_t[(4 * r + lg) * NqNq + j * Nq + li] = acc[r];
  }
  __syncthreads();

  // phase K: k = 4w..4w+3
  double papAcc = 0;  // PAP: sum of q P + qt Gqt over this thread's nodes
#pragma unroll 1
  for (int kk = 0; kk < 4; ++kk) {
    const int k = 4 * w + kk;
    double G[6][4], W[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int g = 0; g < 6; ++g) G[g][r] = Gn[g][r];
      W[r] = Wn[r];
    }
    if (kk < 3) loadG(k + 1);
    d4 qr = {0, 0, 0, 0}, qs = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      // qr = Q D^T : a = q(i=4c+lg, j=li, k), b = D^T[4c+lg][li] = D[li][4c+lg]
      qr = mfma16(s_q[k * NqNq + li * Nq + 4 * c + lg], dA[c], qr);
      // qs = D Q : a = D[li][4c+lg], b = q(i=li, j=4c+lg, k)
      qs = mfma16(dA[c], s_q[k * NqNq + (4 * c + lg) * Nq + li], qs);
    }
    double Gqr[4], Gqs[4], lam[4], qk[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = k * NqNq + (4 * r + lg) * Nq + li;
      const double qt = s_t[n];
      const double qn = s_q[n];
      Gqr[r] = G[0][r] * qr[r] + G[1][r] * qs[r] + G[2][r] * qt;
      Gqs[r] = G[1][r] * qr[r] + G[3][r] * qs[r] + G[4][r] * qt;
      const double Gqt = G[2][r] * qr[r] + G[4][r] * qs[r] + G[5][r] * qt;
      s_t[n] = Gqt;  // in place
      lam[r] = W[r] * lambda * qn;
      s_x[w][(4 * r + lg) * Nq + li] = Gqr[r];
      if (PAP) {
        qk[r] = qn;
        papAcc += qt * Gqt;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    d4 acc = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      // s-part: sum_m D[m][j] Gqs(i,m): a = D[4c+lg][li], b = Gqs(i=li, j=4c+lg) (own register)
      acc = mfma16(dT[c], Gqs[c], acc);
      // r-part: sum_m Gqr(m,j) D[m][i]: a = Gqr(i=4c+lg, j=li) (transposed via LDS), b = D[4c+lg][li]
      acc = mfma16(s_x[w][li * Nq + 4 * c + lg], dT[c], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double P = acc[r] + lam[r];
      s_q[k * NqNq + (4 * r + lg) * Nq + li] = P;
      if (PAP) papAcc += qk[r] * P;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();

  // phase T2: Aq(i,j,k) = partial + sum_m D[m][k] Gqt(i,j,m), j = 4w..4w+3
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int j = 4 * w + jj;
    d4 acc = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = mfma16(dT[c], s_t[(4 * c + lg) * NqNq + j * Nq + li], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = (4 * r + lg) * NqNq + j * Nq + li;
      __builtin_nontemporal_store(acc[r] + s_q[n], Aq + eb + n);
    }
  }
  if constexpr (PAP) {
    __shared__ double s_w[16];
    const double b = block_sum(papAcc, s_w);
    if (t == 0) pap[blockIdx.x] = b;
  }
}

// the PAP form, one partial per element; ax_hex.hip launches the finisher
int launch_ax_mfma_n15_pap(int Nelements, const int* elementList, const int* G2L, const double* wJ,
                           const double* ggeo, const double* D, double lambda, const double* p, double* AqL,
                           double* part, const double* scal, hipStream_t st) {
  hipLaunchKernelGGL(ax_hex_mfma_n15_kernel<true>, dim3(Nelements), dim3(256), 0, st, elementList, G2L, wJ, ggeo, D,
                     lambda, p, AqL, part, scal);
  return launch_status("pcg_ax_hex_pap (N=15)");
}

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_ax_hex_partial_mfma_n15(int Nelements, const int* elementList, const int* GlobalToLocal,
                                            const double* wJ, const double* ggeo, const double* D, double lambda,
                                            const double* q, double* AqL, void* stream) {
  if (Nelements <= 0) return LIBP_OK;
  if (!GlobalToLocal || !wJ || !ggeo || !D || !q || !AqL) return arg_error("ax_hex_mfma_n15", "null pointer");
  hipLaunchKernelGGL(ax_hex_mfma_n15_kernel<false>, dim3(Nelements), dim3(256), 0, as_stream(stream), elementList,
                     GlobalToLocal, wJ, ggeo, D, lambda, q, AqL, nullptr, nullptr);
  return launch_status("ax_hex_mfma_n15");
}
