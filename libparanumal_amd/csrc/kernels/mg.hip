// p-multigrid intergrid transfers for hex elements (FP64 and FP32).
//
// One workgroup per element with an NqF x NqF thread tile; the three 1-D
// contractions with P (NqF x NqC) run as successive sweeps through LDS, as in
// okl/ellipticPreconCoarsenHex3D.okl:209-294 and
// okl/ellipticPreconProlongateHex3D.okl:216-301. The reference uses NqF^3
// threads, which exceeds the 1024-lane workgroup limit for N >= 10; the 2-D
// tile looping over the third index covers every degree up to 15. NQC > 0
// fixes the coarse size at compile time (the 7->3 and 3->1 steps of the
// default schedule), so the sweeps unroll; NQC = 0 takes it at run time. The
// arithmetic and its order are the same either way.
#include "kcommon.hpp"

namespace libp_amd {

template <int NqF, int NQC, typename T>
__global__ void __launch_bounds__(NqF * NqF) mg_coarsen_kernel(const int NqC_, const int* __restrict__ elementList,
                                                               const int* __restrict__ G2L, const T* __restrict__ P,
                                                               const T* __restrict__ qf, const T* __restrict__ w,
                                                               T* __restrict__ qc) {
  constexpr int NpF = NqF * NqF * NqF;
  // LDS: s_f is dead after the k sweep, so the j sweep writes into it (s_ccf).
  // Sizes keep NqF=16 in FP64 within 64 KiB.
  __shared__ T s_f[NqF][NqF][NqF];
  __shared__ T s_cff[NqF - 1][NqF][NqF];
  __shared__ T s_P[NqF][NqF - 1];
  T(*s_ccf)[NqF][NqF] = s_f;
  const int NqC = NQC > 0 ? NQC : NqC_;
  const int e = blockIdx.x;
  const int element = elementList ? elementList[e] : e;
  const int t = threadIdx.x, i = t % NqF, j = t / NqF;
  const int NpC = NqC * NqC * NqC;
  // index loads first, then unconditional value loads (masked nodes read entry 0 and take 0):
  // loads under (id >= 0) compile to branches that wait for each load in turn
  int ids[NqF];
#pragma unroll
  for (int k = 0; k < NqF; ++k) ids[k] = G2L[(size_t)element * NpF + k * NqF * NqF + t];
#pragma unroll
  for (int k = 0; k < NqF; ++k) {
    const int id = max(ids[k], 0);
    // w: the inverse-multiplicity weight of the fine residual (ellipticPreconMultiGridLevel.cpp:51),
    // applied here instead of in a separate pass (the same single rounding)
    const T qv = qf[id];
    const T wv = w ? w[id] : T(1);
    s_f[k][j][i] = (ids[k] >= 0) ? (w ? T(1) * wv * qv : qv) : T(0);
  }
  if (t < NqF * NqC) s_P[t / NqC][t % NqC] = P[t];
  __syncthreads();
  for (int k = 0; k < NqC; ++k) {
    T res = 0;
    for (int m = 0; m < NqF; ++m) res += s_P[m][k] * s_f[m][j][i];
    s_cff[k][j][i] = res;
  }
  __syncthreads();
  if (j < NqC)
    for (int k = 0; k < NqC; ++k) {
      T res = 0;
      for (int m = 0; m < NqF; ++m) res += s_P[m][j] * s_cff[k][m][i];
      s_ccf[k][j][i] = res;
    }
  __syncthreads();
  if (i < NqC && j < NqC) {
    for (int k = 0; k < NqC; ++k) {
      T res = 0;
      for (int m = 0; m < NqF; ++m) res += s_P[m][i] * s_ccf[k][j][m];
      qc[(size_t)element * NpC + i + j * NqC + k * NqC * NqC] = res;
    }
  }
}

template <int NqF, int NQC, typename T>
__global__ void __launch_bounds__(NqF * NqF) mg_prolong_kernel(const int NqC_, const int* __restrict__ elementList,
                                                               const int* __restrict__ G2LC, const T* __restrict__ P,
                                                               const T* __restrict__ qcIn, T* __restrict__ qN) {
  constexpr int NpF = NqF * NqF * NqF;
  // LDS: s_ccc is dead after the i sweep, so the j sweep writes into it (s_cff).
  __shared__ T s_A[NqF - 1][NqF][NqF];
  __shared__ T s_ccf[NqF - 1][NqF - 1][NqF];
  __shared__ T s_P[NqF][NqF - 1];
  T(*s_ccc)[NqF][NqF] = s_A;
  T(*s_cff)[NqF][NqF] = s_A;
  const int NqC = NQC > 0 ? NQC : NqC_;
  const int e = blockIdx.x;
  const int element = elementList ? elementList[e] : e;
  const int t = threadIdx.x, i = t % NqF, j = t / NqF;
  const int NpC = NqC * NqC * NqC;
  if (i < NqC && j < NqC) {
    // index loads first, then unconditional value loads (a masked node reads entry 0 and
    // takes 0); NqC < NqF, and k < NqC is the same for every lane
    int ids[NqF - 1];
#pragma unroll
    for (int k = 0; k < NqF - 1; ++k)
      if (k < NqC) ids[k] = G2LC[(size_t)element * NpC + i + j * NqC + k * NqC * NqC];
#pragma unroll
    for (int k = 0; k < NqF - 1; ++k)
      if (k < NqC) {
        const T qv = qcIn[max(ids[k], 0)];
        s_ccc[k][j][i] = (ids[k] >= 0) ? qv : T(0);
      }
  }
  if (t < NqF * NqC) s_P[t / NqC][t % NqC] = P[t];
  __syncthreads();
  if (j < NqC)
    for (int k = 0; k < NqC; ++k) {
      T res = 0;
      for (int m = 0; m < NqC; ++m) res += s_P[i][m] * s_ccc[k][j][m];
      s_ccf[k][j][i] = res;
    }
  __syncthreads();
  for (int k = 0; k < NqC; ++k) {
    T res = 0;
    for (int m = 0; m < NqC; ++m) res += s_P[j][m] * s_ccf[k][m][i];
    s_cff[k][j][i] = res;
  }
  __syncthreads();
  for (int k = 0; k < NqF; ++k) {
    T res = 0;
    for (int m = 0; m < NqC; ++m) res += s_P[k][m] * s_cff[m][j][i];
    qN[(size_t)element * NpF + i + j * NqF + k * NqF * NqF] = res;
  }
}

template <typename T, bool COARSEN>
static int transfer(int Nf, int Nc, int Nelements, const int* elist, const int* g2l, const T* P, const T* in, T* out,
                    void* stream, const T* w = nullptr) {
  if (Nelements <= 0) return LIBP_OK;
  if (Nc < 1 || Nc >= Nf || Nf > 15) return arg_error("mg_transfer", "need 1 <= Nc < Nf <= 15");
  if (!g2l || !P || !in || !out) return arg_error("mg_transfer", "null pointer");
  hipStream_t st = as_stream(stream);
  const int NqC = Nc + 1;
  // the default schedule's steps with the coarse size fixed at compile time
  if (Nf == 7 && Nc == 3) {
    if (COARSEN)
      hipLaunchKernelGGL((mg_coarsen_kernel<8, 4, T>), dim3(Nelements), dim3(64), 0, st, NqC, elist, g2l, P, in, w,
                         out);
    else
      hipLaunchKernelGGL((mg_prolong_kernel<8, 4, T>), dim3(Nelements), dim3(64), 0, st, NqC, elist, g2l, P, in, out);
    return launch_status(COARSEN ? "mg_coarsen" : "mg_prolong");
  }
  if (Nf == 3 && Nc == 1) {
    if (COARSEN)
      hipLaunchKernelGGL((mg_coarsen_kernel<4, 2, T>), dim3(Nelements), dim3(16), 0, st, NqC, elist, g2l, P, in, w,
                         out);
    else
      hipLaunchKernelGGL((mg_prolong_kernel<4, 2, T>), dim3(Nelements), dim3(16), 0, st, NqC, elist, g2l, P, in, out);
    return launch_status(COARSEN ? "mg_coarsen" : "mg_prolong");
  }
  switch (Nf + 1) {
#define LIBP_MG_CASE(NQ)                                                                                            \
  case NQ:                                                                                                          \
    if (COARSEN)                                                                                                    \
      hipLaunchKernelGGL((mg_coarsen_kernel<NQ, 0, T>), dim3(Nelements), dim3(NQ * NQ), 0, st, NqC, elist, g2l, P, \
                         in, w, out);                                                                               \
    else                                                                                                            \
      hipLaunchKernelGGL((mg_prolong_kernel<NQ, 0, T>), dim3(Nelements), dim3(NQ * NQ), 0, st, NqC, elist, g2l, P,  \
                         in, out);                                                                                  \
    break;
    LIBP_MG_CASE(3)
    LIBP_MG_CASE(4)
    LIBP_MG_CASE(5)
    LIBP_MG_CASE(6)
    LIBP_MG_CASE(7)
    LIBP_MG_CASE(8)
    LIBP_MG_CASE(9)
    LIBP_MG_CASE(10)
    LIBP_MG_CASE(11)
    LIBP_MG_CASE(12)
    LIBP_MG_CASE(13)
    LIBP_MG_CASE(14)
    LIBP_MG_CASE(15)
    LIBP_MG_CASE(16)
#undef LIBP_MG_CASE
    default:
      return arg_error("mg_transfer", "unsupported fine degree");
  }
  return launch_status(COARSEN ? "mg_coarsen" : "mg_prolong");
}

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_mg_coarsen_hex_f64(int Nf, int Nc, int Nelements, const int* elementList,
                                       const int* GlobalToLocal, const double* P, const double* qf, double* qc,
                                       void* stream) {
  return transfer<double, true>(Nf, Nc, Nelements, elementList, GlobalToLocal, P, qf, qc, stream);
}
extern "C" int libp_mg_coarsen_hex_f32(int Nf, int Nc, int Nelements, const int* elementList,
                                       const int* GlobalToLocal, const float* P, const float* qf, float* qc,
                                       void* stream) {
  return transfer<float, true>(Nf, Nc, Nelements, elementList, GlobalToLocal, P, qf, qc, stream);
}
extern "C" int libp_mg_prolong_hex_f64(int Nf, int Nc, int Nelements, const int* elementList,
                                       const int* GlobalToLocalC, const double* P, const double* qc, double* qf,
                                       void* stream) {
  return transfer<double, false>(Nf, Nc, Nelements, elementList, GlobalToLocalC, P, qc, qf, stream);
}
extern "C" int libp_mg_prolong_hex_f32(int Nf, int Nc, int Nelements, const int* elementList,
                                       const int* GlobalToLocalC, const float* P, const float* qc, float* qf,
                                       void* stream) {
  return transfer<float, false>(Nf, Nc, Nelements, elementList, GlobalToLocalC, P, qc, qf, stream);
}

extern "C" int libp_mg_coarsen_weighted_hex_f64(int Nf, int Nc, int Nelements, const int* elementList,
                                                const int* GlobalToLocal, const double* P, const double* weight,
                                                const double* qf, double* qc, void* stream) {
  if (!weight) return arg_error("mg_coarsen_weighted", "null pointer");
  return transfer<double, true>(Nf, Nc, Nelements, elementList, GlobalToLocal, P, qf, qc, stream, weight);
}
extern "C" int libp_mg_coarsen_weighted_hex_f32(int Nf, int Nc, int Nelements, const int* elementList,
                                                const int* GlobalToLocal, const float* P, const float* weight,
                                                const float* qf, float* qc, void* stream) {
  if (!weight) return arg_error("mg_coarsen_weighted", "null pointer");
  return transfer<float, true>(Nf, Nc, Nelements, elementList, GlobalToLocal, P, qf, qc, stream, weight);
}
