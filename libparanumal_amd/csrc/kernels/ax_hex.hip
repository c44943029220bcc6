// Matrix-free C0 hex stiffness + mass operator for gfx950.
//
// Thread layout: one Nq x Nq thread tile per element, each thread owning the
// k-pencil q(i,j,0:N) in registers (the structure of
// solvers/elliptic/okl/ellipticAxHex3D.okl:156-295). Differences that matter on
// CDNA4:
//  * several small elements share a 64-lane wavefront (NE elements per block),
//    so N<=4 levels do not idle most of the lanes;
//  * the geometric factors, wJ and GlobalToLocal are each read exactly once,
//    so they are loaded non-temporally and leave L2 to the gathered q values
//    that neighbouring elements share (+9% on the 48^3, N=7 box);
//  * an explicit waves-per-EU target bounds the register budget, keeping
//    enough wavefronts resident to cover HBM latency without spilling; at
//    N=7 four elements (four wavefronts) per workgroup at 4 waves/SIMD was
//    the fastest of the measured shapes (+4% over one element per group);
//  * elements run in natural order: brick or XCD-contiguous orderings that cut
//    q re-fetch were all slower (they break the sequential page streams of
//    the dominant arrays);
//  * the E-vector result is written with non-temporal stores (it is consumed
//    by the gather kernel, not re-read by this one).
//  * lambda = 0 (the Poisson operator) runs an instance that does not read wJ
//    (the reference reads it whatever lambda is): the mass term is +-0 in both,
//    so the result is the same to the bit, and the FP64 kernel streams 60
//    instead of 68 bytes per local node (48^3, N=7: 649 -> 604 us). Testing
//    lambda inside the k-loop instead gave the same bits at 2.4x the time: the
//    branches split the loop into blocks the scheduler no longer overlapped.
//
// The split output form (OUT = 1) serves the whole C0 operator A = Q^T A_L Q
// (ellipticOperator.cpp:29-107) without writing the full E-vector: a node
// interior to its element has exactly one copy, so its gathered row is the
// element's value itself and is stored straight into the gathered output
// (0 + a = a: bit-identical to the gather). Only the element-boundary ("face")
// nodes go to a compact face E-vector, e * Nface + faceIndex(n), which a gather
// over the shared rows alone then sums (ogs_gather_face_kernel). The ids of
// the element's nodes wait in LDS for the output stage (the kernel has no
// registers to spare at N=7). With PAP, the block also forms its share of
// p.Ap over the interior rows from the q and Aq values it holds.
#include "reduce.hpp"

namespace libp_amd {

template <int Nq, typename T = double>
struct AxShape {
  static constexpr bool f32 = sizeof(T) == 4;
  static constexpr int NqNq = Nq * Nq;
  static constexpr int Np = Nq * Nq * Nq;
  // elements per workgroup: fill a wavefront for small N. At N=7 the FP64
  // kernel measured fastest with four one-wavefront elements per workgroup,
  // the FP32 kernel with two (32^3 box: 101.1 us, against 105.0 us with one
  // element and eight waves per SIMD, 101.4 us with four, 211 us with two and
  // eight waves).
  static constexpr int NE = (Nq == 8) ? (f32 ? 2 : 4) : ((NqNq >= 64) ? 1 : (64 / NqNq));
  static constexpr int threads = NE * NqNq;
  // occupancy target: bounds the register budget so that the compiler
  // re-reads loop-invariant LDS operands instead of pinning them in VGPRs
  static constexpr int waves = (Nq <= 8) ? 4 : 2;
};

// Number of element-boundary nodes of an Nq^3 element, and the position of
// node (i,j,k) among them in local order (the node index minus the interior
// nodes before it); -1 for an interior node.
__host__ __device__ constexpr int hex_face_nodes(int Nq) { return Nq * Nq * Nq - (Nq - 2) * (Nq - 2) * (Nq - 2); }
__device__ inline int interior_before(int x, int Nq) { return min(max(x - 1, 0), Nq - 2); }
__device__ inline bool hex_interior(int i, int j, int k, int Nq) {
  return i > 0 && i < Nq - 1 && j > 0 && j < Nq - 1 && k > 0 && k < Nq - 1;
}
__device__ inline int hex_face_index(int i, int j, int k, int Nq) {
  const int n = i + j * Nq + k * Nq * Nq;
  const bool kin = k > 0 && k < Nq - 1, jin = j > 0 && j < Nq - 1;
  const int before = interior_before(k, Nq) * (Nq - 2) * (Nq - 2) +
                     (kin ? interior_before(j, Nq) * (Nq - 2) + (jin ? interior_before(i, Nq) : 0) : 0);
  return n - before;
}

// Block sum of one double per thread into part[blockIdx.x] (thread 0 writes).
template <int threads>
__device__ inline void block_partial(const double acc, double* __restrict__ part) {
  const int t = threadIdx.x;
  if constexpr (threads % 64 == 0) {
    __shared__ double s_w[threads / 64];
    const double b = block_sum(acc, s_w);
    if (t == 0) part[blockIdx.x] = b;
  } else {  // partial wavefront (N <= 4): a tree in LDS, no cross-lane reads of idle lanes
    constexpr int P = (threads <= 64) ? 64 : 256;
    __shared__ double s_red[P];
    s_red[t] = acc;
    for (int z = threads + t; z < P; z += threads) s_red[z] = 0.0;
#pragma unroll
    for (int off = P / 2; off > 0; off >>= 1) {
      __syncthreads();
      if (t < off) s_red[t] += s_red[t + off];
    }
    if (t == 0) part[blockIdx.x] = s_red[0];
  }
}

// MODE 0: q is a gathered vector indexed through GlobalToLocal (-1 -> 0).
// MODE 1: q is an E-vector (element-local input, no gather).
// OUT 0: Aq is the E-vector; with PAP (MODE 0, the PCG form), each block
// also stores sum over its local nodes of q_L * (A_L q)_L, which summed over
// all blocks is p.Ap = p^T Q^T A_L Q p for a gathered q = p, and the kernel
// returns at once while scal[8] (STOP) is set. OUT 1 (MODE 0 only): interior nodes go to the
// gathered Ag, element-boundary nodes to the face E-vector Aq; PAP adds one
// partial of sum q Ag over the interior rows per block, pap[pap0 + block].
// Every element slot of a block is a real element: the launcher runs the
// whole blocks first and the remainder with NE=1 (element offset e0). A
// clamped/predicated tail inside this kernel defeats the compiler's uniform
// address analysis and roughly quadruples register use.
template <int Nq, int NE, typename T, int MODE, int OUT = 0, bool PAP = false, bool LAM = true>
__global__ void __launch_bounds__(NE * Nq * Nq) __attribute__((amdgpu_waves_per_eu(AxShape<Nq, T>::waves, 8)))
    ax_hex_kernel(const int e0, const int* __restrict__ elementList, const int* __restrict__ G2L,
                  const T* __restrict__ wJ, const T* __restrict__ ggeo, const T* __restrict__ D, const T lambda,
                  const T* __restrict__ q, T* __restrict__ Aq, T* __restrict__ Ag = nullptr,
                  double* __restrict__ pap = nullptr, const int pap0 = 0,
                  const int* __restrict__ fslot = nullptr, const double* __restrict__ scal = nullptr) {
  // a*b + c contracts only within one source expression, so every
  // instantiation (E-vector, split, split with p.Ap) rounds alike: with
  // contraction across expressions the backend may fuse x*y + z*w as
  // fma(x, y, z*w) in one instantiation and fma(z, w, x*y) in another
#pragma clang fp contract(on)
  using S = AxShape<Nq, T>;
  constexpr int NqNq = S::NqNq, Np = S::Np, threads = NE * NqNq;
  __shared__ T s_D[Nq][Nq + 1];
  __shared__ T s_q[NE][Nq][Nq + 1];
  __shared__ T s_Gqr[NE][Nq][Nq + 1];
  __shared__ T s_Gqs[NE][Nq][Nq + 1];
  __shared__ int s_id[OUT == 1 ? NE : 1][OUT == 1 ? Nq : 1][OUT == 1 ? NqNq : 1];
  __shared__ T s_out[OUT == 1 ? NE : 1][OUT == 1 ? Nq : 1][OUT == 1 ? NqNq : 1];

  if (scal != nullptr && scal[8] != 0.0) return;  // a PCG step after convergence (STOP)
  const int t = threadIdx.x;
  const int es = (NE == 1) ? 0 : t / NqNq;
  const int ij = (NE == 1) ? t : t - es * NqNq;
  const int i = ij % Nq;
  const int j = ij / Nq;
  const int eIdx = e0 + blockIdx.x * NE + es;
  const int element = elementList ? elementList[eIdx] : eIdx;

  for (int n = t; n < NqNq; n += threads) s_D[n / Nq][n % Nq] = D[n];

  T r_q[Nq], r_Aq[Nq];
  const size_t base = (size_t)element * Np + ij;
  if (MODE == 0) {
    // all Nq index loads in flight, then all Nq value loads: the value loads are
    // unconditional (a masked node, id -1, reads q[0] and takes 0), since a load
    // under (id >= 0) compiles to a branch that waits for each load in turn
    int ids[Nq];
#pragma unroll
    for (int k = 0; k < Nq; ++k) ids[k] = __builtin_nontemporal_load(G2L + base + k * NqNq);
#pragma unroll
    for (int k = 0; k < Nq; ++k) {
      const T v = q[max(ids[k], 0)];
      r_q[k] = (ids[k] >= 0) ? v : T(0);
      if (OUT == 1) s_id[es][k][ij] = ids[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < Nq; ++k) r_q[k] = __builtin_nontemporal_load(q + base + k * NqNq);
  }
#pragma unroll
  for (int k = 0; k < Nq; ++k) r_Aq[k] = 0;
  const T* gb = ggeo + (size_t)element * 6 * Np + ij;
  const T* wb = wJ + (size_t)element * Np + ij;
  __syncthreads();

#pragma unroll
  for (int k = 0; k < Nq; ++k) {
    const T G00 = __builtin_nontemporal_load(gb + 0 * Np + k * NqNq);
    const T G01 = __builtin_nontemporal_load(gb + 1 * Np + k * NqNq);
    const T G02 = __builtin_nontemporal_load(gb + 2 * Np + k * NqNq);
    const T G11 = __builtin_nontemporal_load(gb + 3 * Np + k * NqNq);
    const T G12 = __builtin_nontemporal_load(gb + 4 * Np + k * NqNq);
    const T G22 = __builtin_nontemporal_load(gb + 5 * Np + k * NqNq);
    // LAM = false (lambda = 0): wJ is not read; lambda * GwJ * q is +-0 with either GwJ
    const T GwJ = LAM ? __builtin_nontemporal_load(wb + k * NqNq) : T(0);

    s_q[es][j][i] = r_q[k];
    T qt = 0;
#pragma unroll
    for (int m = 0; m < Nq; ++m) qt += s_D[k][m] * r_q[m];
    __syncthreads();

    T qr = 0, qs = 0;
#pragma unroll
    for (int m = 0; m < Nq; ++m) {
      qr += s_D[i][m] * s_q[es][j][m];
      qs += s_D[j][m] * s_q[es][m][i];
    }
    s_Gqs[es][j][i] = G01 * qr + G11 * qs + G12 * qt;
    s_Gqr[es][j][i] = G00 * qr + G01 * qs + G02 * qt;
    const T Gqt = G02 * qr + G12 * qs + G22 * qt;
    T Auk = GwJ * lambda * r_q[k];
    __syncthreads();

#pragma unroll
    for (int m = 0; m < Nq; ++m) {
      Auk += s_D[m][j] * s_Gqs[es][m][i];
      r_Aq[m] += s_D[k][m] * Gqt;
      Auk += s_D[m][i] * s_Gqr[es][j][m];
    }
    r_Aq[k] += Auk;
  }

  if (OUT == 0) {
    T* out = Aq + base;
#pragma unroll
    for (int k = 0; k < Nq; ++k) __builtin_nontemporal_store(r_Aq[k], out + k * NqNq);
    if constexpr (PAP) {
      double acc = 0;
#pragma unroll
      for (int k = 0; k < Nq; ++k) acc += (double)r_q[k] * (double)r_Aq[k];
      block_partial<threads>(acc, pap + pap0);
    }
  } else {
    // The results go through LDS first: stored with the regular pattern of the
    // E-vector store, the registers die there. Scattering straight from the
    // registers (one branch and address per k) made the compiler spill ~500
    // VGPRs at N=7.
    constexpr int Nface = hex_face_nodes(Nq);
    double acc = 0;
#pragma unroll
    for (int k = 0; k < Nq; ++k) {
      s_out[es][k][ij] = r_Aq[k];
      if (PAP) acc += hex_interior(i, j, k, Nq) ? (double)r_q[k] * (double)r_Aq[k] : 0.0;
    }
    __syncthreads();
    T* fout = Aq + (size_t)element * Nface;
    for (int k = 0; k < Nq; ++k) {
      const T v = s_out[es][k][ij];
      if (hex_interior(i, j, k, Nq)) {
        const int id = s_id[es][k][ij];
        if (id >= 0) Ag[id] = v;
      } else {
        const int fi = hex_face_index(i, j, k, Nq);
        if (fslot)  // slot layout: the copy's place in the row-ordered copy list
          Aq[__builtin_nontemporal_load(fslot + (size_t)element * Nface + fi)] = v;
        else
          __builtin_nontemporal_store(v, fout + fi);
      }
    }
    if constexpr (PAP) block_partial<threads>(acc, pap + pap0);
  }
}

// lambda = 0 (the Poisson operator) takes the instances that do not read wJ:
// the mass term lambda * wJ * q is +-0 in both, so the results are the same to
// the bit, and the kernel moves 8 (FP64) or 4 (FP32) bytes less per local node.
template <int Nq, typename T, int MODE, bool LAM>
static void launch_ax_lam(int Nelements, const int* elist, const int* g2l, const T* wJ, const T* ggeo, const T* D,
                          T lambda, const T* q, T* Aq, hipStream_t st) {
  using S = AxShape<Nq, T>;
  const int full = Nelements / S::NE;
  if (full > 0)
    hipLaunchKernelGGL((ax_hex_kernel<Nq, S::NE, T, MODE, 0, false, LAM>), dim3(full), dim3(S::threads), 0, st, 0,
                       elist, g2l, wJ, ggeo, D, lambda, q, Aq);
  const int rest = Nelements - full * S::NE;
  if (S::NE > 1 && rest > 0)
    hipLaunchKernelGGL((ax_hex_kernel<Nq, 1, T, MODE, 0, false, LAM>), dim3(rest), dim3(S::NqNq), 0, st,
                       full * S::NE, elist, g2l, wJ, ggeo, D, lambda, q, Aq);
}
template <int Nq, typename T, int MODE>
static int launch_ax(int Nelements, const int* elist, const int* g2l, const T* wJ, const T* ggeo, const T* D,
                     T lambda, const T* q, T* Aq, hipStream_t st) {
  if (Nelements <= 0) return LIBP_OK;
  if (lambda == T(0))
    launch_ax_lam<Nq, T, MODE, false>(Nelements, elist, g2l, wJ, ggeo, D, lambda, q, Aq, st);
  else
    launch_ax_lam<Nq, T, MODE, true>(Nelements, elist, g2l, wJ, ggeo, D, lambda, q, Aq, st);
  return launch_status("ax_hex");
}

// Split output (OUT = 1): interior rows into Ag, boundary nodes into the face
// E-vector AqF, and with pap one p.Ap partial per block (split_blocks of them).
template <int Nq>
static int split_blocks(int Nelements) {
  using S = AxShape<Nq, double>;
  return Nelements / S::NE + ((S::NE > 1) ? Nelements % S::NE : 0);
}

template <int Nq, bool LAM>
static int launch_ax_split_lam(int Nelements, const int* elist, const int* g2l, const double* wJ, const double* ggeo,
                           const double* D, double lambda, const double* q, double* AqF, double* Ag, double* pap,
                           const int* fslot, hipStream_t st) {
  using S = AxShape<Nq, double>;
  if (Nelements <= 0) return LIBP_OK;
  const int full = Nelements / S::NE;
  const int rest = Nelements - full * S::NE;
  if (pap) {
    if (full > 0)
      hipLaunchKernelGGL((ax_hex_kernel<Nq, S::NE, double, 0, 1, true, LAM>), dim3(full), dim3(S::threads), 0, st, 0,
                         elist, g2l, wJ, ggeo, D, lambda, q, AqF, Ag, pap, 0, fslot);
    if (S::NE > 1 && rest > 0)
      hipLaunchKernelGGL((ax_hex_kernel<Nq, 1, double, 0, 1, true, LAM>), dim3(rest), dim3(S::NqNq), 0, st,
                         full * S::NE, elist, g2l, wJ, ggeo, D, lambda, q, AqF, Ag, pap, full, fslot);
  } else {
    if (full > 0)
      hipLaunchKernelGGL((ax_hex_kernel<Nq, S::NE, double, 0, 1, false, LAM>), dim3(full), dim3(S::threads), 0, st, 0,
                         elist, g2l, wJ, ggeo, D, lambda, q, AqF, Ag, nullptr, 0, fslot);
    if (S::NE > 1 && rest > 0)
      hipLaunchKernelGGL((ax_hex_kernel<Nq, 1, double, 0, 1, false, LAM>), dim3(rest), dim3(S::NqNq), 0, st,
                         full * S::NE, elist, g2l, wJ, ggeo, D, lambda, q, AqF, Ag, nullptr, 0, fslot);
  }
  return launch_status("ax_hex_split");
}
template <int Nq>
static int launch_ax_split(int Nelements, const int* elist, const int* g2l, const double* wJ, const double* ggeo,
                           const double* D, double lambda, const double* q, double* AqF, double* Ag, double* pap,
                           const int* fslot, hipStream_t st) {
  if (lambda == 0.0)
    return launch_ax_split_lam<Nq, false>(Nelements, elist, g2l, wJ, ggeo, D, lambda, q, AqF, Ag, pap, fslot, st);
  return launch_ax_split_lam<Nq, true>(Nelements, elist, g2l, wJ, ggeo, D, lambda, q, AqF, Ag, pap, fslot, st);
}

// p.Ap from the partials of the PCG-form Ax kernel, summed in block order,
// then alpha = s[RDOTZ1] / p.Ap (linearSolverPCG.cpp:134-140).
__global__ void __launch_bounds__(1024) ax_pap_finish_kernel(const int nb, double* __restrict__ s,
                                                             const double* __restrict__ part) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  // 27,648 partials at 48^3, N=7: one round of 32 loads per thread (7.3 us with 8)
  const double pap = block_sum(strided_sum<32>(part, nb), s_w);
  if (threadIdx.x == 0) {
    s[4] = pap;
    s[2] = s[0] / pap;
  }
}

// Partial Ax of p (gathered, through GlobalToLocal) into the E-vector with
// p.Ap from the local nodes and alpha (MODE 0, OUT 0, PAP).
template <int Nq, bool LAM>
static int launch_ax_pap_lam(int Nelements, const int* elist, const int* g2l, const double* wJ, const double* ggeo,
                         const double* D, double lambda, const double* p, double* AqL, double* part, double* s,
                         hipStream_t st) {
  using S = AxShape<Nq, double>;
  if (Nelements <= 0) return LIBP_OK;
  const int full = Nelements / S::NE;
  const int rest = Nelements - full * S::NE;
  if (full > 0)
    hipLaunchKernelGGL((ax_hex_kernel<Nq, S::NE, double, 0, 0, true, LAM>), dim3(full), dim3(S::threads), 0, st, 0, elist,
                       g2l, wJ, ggeo, D, lambda, p, AqL, nullptr, part, 0, nullptr, s);
  if (S::NE > 1 && rest > 0)
    hipLaunchKernelGGL((ax_hex_kernel<Nq, 1, double, 0, 0, true, LAM>), dim3(rest), dim3(S::NqNq), 0, st, full * S::NE,
                       elist, g2l, wJ, ggeo, D, lambda, p, AqL, nullptr, part, full, nullptr, s);
  hipLaunchKernelGGL(ax_pap_finish_kernel, dim3(1), dim3(1024), 0, st, split_blocks<Nq>(Nelements), s,
                     (const double*)part);
  return launch_status("pcg_ax_hex_pap");
}
template <int Nq>
static int launch_ax_pap(int Nelements, const int* elist, const int* g2l, const double* wJ, const double* ggeo,
                         const double* D, double lambda, const double* p, double* AqL, double* part, double* s,
                         hipStream_t st) {
  if (lambda == 0.0) return launch_ax_pap_lam<Nq, false>(Nelements, elist, g2l, wJ, ggeo, D, lambda, p, AqL, part, s, st);
  return launch_ax_pap_lam<Nq, true>(Nelements, elist, g2l, wJ, ggeo, D, lambda, p, AqL, part, s, st);
}

template <typename T, int MODE>
static int dispatch_ax(int N, int Nelements, const int* elist, const int* g2l, const T* wJ, const T* ggeo,
                       const T* D, T lambda, const T* q, T* Aq, void* stream) {
  hipStream_t st = as_stream(stream);
  if (!wJ || !ggeo || !D || !q || !Aq || (MODE == 0 && !g2l)) return arg_error("ax_hex", "null pointer");
  switch (N) {
#define LIBP_AX_CASE(NN) \
  case NN:               \
    return launch_ax<NN + 1, T, MODE>(Nelements, elist, g2l, wJ, ggeo, D, lambda, q, Aq, st);
    LIBP_AX_CASE(1)
    LIBP_AX_CASE(2)
    LIBP_AX_CASE(3)
    LIBP_AX_CASE(4)
    LIBP_AX_CASE(5)
    LIBP_AX_CASE(6)
    LIBP_AX_CASE(7)
    LIBP_AX_CASE(8)
    LIBP_AX_CASE(9)
    LIBP_AX_CASE(10)
    LIBP_AX_CASE(11)
    LIBP_AX_CASE(12)
    LIBP_AX_CASE(13)
    LIBP_AX_CASE(14)
    LIBP_AX_CASE(15)
#undef LIBP_AX_CASE
    default:
      return arg_error("ax_hex", "polynomial degree must be in [1,15]");
  }
}

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_ax_hex_partial_f64(int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                                       const double* wJ, const double* ggeo, const double* D, double lambda,
                                       const double* q, double* AqL, void* stream) {
  return dispatch_ax<double, 0>(N, Nelements, elementList, GlobalToLocal, wJ, ggeo, D, lambda, q, AqL, stream);
}

extern "C" int libp_ax_hex_partial_f32(int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                                       const float* wJ, const float* ggeo, const float* D, float lambda,
                                       const float* q, float* AqL, void* stream) {
  return dispatch_ax<float, 0>(N, Nelements, elementList, GlobalToLocal, wJ, ggeo, D, lambda, q, AqL, stream);
}

extern "C" int libp_ax_hex_local_f64(int N, int Nelements, const double* wJ, const double* ggeo, const double* D,
                                     double lambda, const double* qL, double* AqL, void* stream) {
  return dispatch_ax<double, 1>(N, Nelements, nullptr, nullptr, wJ, ggeo, D, lambda, qL, AqL, stream);
}

extern "C" int libp_ax_hex_split_blocks(int N, int Nelements) {
  switch (N) {
#define LIBP_AX_CASE(NN) \
  case NN:               \
    return split_blocks<NN + 1>(Nelements);
    LIBP_AX_CASE(1)
    LIBP_AX_CASE(2)
    LIBP_AX_CASE(3)
    LIBP_AX_CASE(4)
    LIBP_AX_CASE(5)
    LIBP_AX_CASE(6)
    LIBP_AX_CASE(7)
    LIBP_AX_CASE(8)
    LIBP_AX_CASE(9)
    LIBP_AX_CASE(10)
    LIBP_AX_CASE(11)
    LIBP_AX_CASE(12)
    LIBP_AX_CASE(13)
    LIBP_AX_CASE(14)
#undef LIBP_AX_CASE
    case 15:  // the MFMA kernel of the PCG form: one partial per element
      return Nelements;
    default:
      return -1;
  }
}

extern "C" int libp_ax_hex_face_nodes(int N) { return (N >= 1 && N <= 15) ? hex_face_nodes(N + 1) : -1; }

extern "C" int libp_ax_hex_split_f64(int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                                     const double* wJ, const double* ggeo, const double* D, double lambda,
                                     const double* q, double* Aq, double* AqFace, const int* faceSlots,
                                     double* papPartials, void* stream) {
  if (!GlobalToLocal || !wJ || !ggeo || !D || !q || !Aq || !AqFace) return arg_error("ax_hex_split", "null pointer");
  hipStream_t st = as_stream(stream);
  switch (N) {
#define LIBP_AX_CASE(NN) \
  case NN:               \
    return launch_ax_split<NN + 1>(Nelements, elementList, GlobalToLocal, wJ, ggeo, D, lambda, q, AqFace, Aq, \
                                   papPartials, faceSlots, st);
    LIBP_AX_CASE(1)
    LIBP_AX_CASE(2)
    LIBP_AX_CASE(3)
    LIBP_AX_CASE(4)
    LIBP_AX_CASE(5)
    LIBP_AX_CASE(6)
    LIBP_AX_CASE(7)
    LIBP_AX_CASE(8)
    LIBP_AX_CASE(9)
    LIBP_AX_CASE(10)
    LIBP_AX_CASE(11)
    LIBP_AX_CASE(12)
    LIBP_AX_CASE(13)
    LIBP_AX_CASE(14)
#undef LIBP_AX_CASE
    default:
      return arg_error("ax_hex_split", "polynomial degree must be in [1,14]");
  }
}

namespace libp_amd {
// ax_hex_mfma.hip: the N=15 MFMA kernel in its PCG form (p.Ap partials, STOP)
int launch_ax_mfma_n15_pap(int Nelements, const int* elementList, const int* G2L, const double* wJ,
                           const double* ggeo, const double* D, double lambda, const double* p, double* AqL,
                           double* part, const double* scal, hipStream_t st);
}  // namespace libp_amd

extern "C" int libp_pcg_ax_hex_pap_f64(int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                                       const double* wJ, const double* ggeo, const double* D, double lambda,
                                       const double* p, double* AqL, double* papPartials, double* scalars,
                                       void* stream) {
  if (!GlobalToLocal || !wJ || !ggeo || !D || !p || !AqL || !papPartials || !scalars)
    return arg_error("pcg_ax_hex_pap", "null pointer");
  hipStream_t st = as_stream(stream);
  switch (N) {
#define LIBP_AX_CASE(NN) \
  case NN:               \
    return launch_ax_pap<NN + 1>(Nelements, elementList, GlobalToLocal, wJ, ggeo, D, lambda, p, AqL, papPartials, \
                                 scalars, st);
    LIBP_AX_CASE(1)
    LIBP_AX_CASE(2)
    LIBP_AX_CASE(3)
    LIBP_AX_CASE(4)
    LIBP_AX_CASE(5)
    LIBP_AX_CASE(6)
    LIBP_AX_CASE(7)
    LIBP_AX_CASE(8)
    LIBP_AX_CASE(9)
    LIBP_AX_CASE(10)
    LIBP_AX_CASE(11)
    LIBP_AX_CASE(12)
    LIBP_AX_CASE(13)
    LIBP_AX_CASE(14)
#undef LIBP_AX_CASE
    case 15: {  // the FP64 MFMA kernel, one element (and one partial) per workgroup
      if (Nelements <= 0) return LIBP_OK;
      const int rc = launch_ax_mfma_n15_pap(Nelements, elementList, GlobalToLocal, wJ, ggeo, D, lambda, p, AqL,
                                            papPartials, scalars, st);
      if (rc != LIBP_OK) return rc;
      hipLaunchKernelGGL(ax_pap_finish_kernel, dim3(1), dim3(1024), 0, st, Nelements, scalars,
                         (const double*)papPartials);
      return launch_status("pcg_ax_hex_pap");
    }
    default:
      return arg_error("pcg_ax_hex_pap", "polynomial degree must be in [1,15]");
  }
}
