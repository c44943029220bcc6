// Single-rank gather-scatter kernels (CSR rows, a thread per gathered row).
//
// The reference stages blocks of <= 512 nonzeros through shared memory
// (libs/ogs/okl/ogsKernels.okl:86-164). On the C0 hex path every row holds
// 1-8 entries (the copies of one node), so a thread per row needs no LDS
// staging and no barrier. A row of up to 8 entries issues all of its column
// loads, then all of its value loads, before summing: the copies of a node
// lie in different elements, and a loop of dependent loads would leave the
// memory system idle. Sums still run from zero in ascending column order,
// which is the reference's order, so Gather results are bit-identical.
#include <algorithm>

#include "gather_block.hpp"
#include "reduce.hpp"

namespace libp_amd {

// The first eight copies of two CSR rows (a[r], len[r]) for the gather
// kernels: the index loads, then the value loads, each up to the longest row
// of the wavefront, all issued before any is used. A load under the lane's
// own `k < len` is compiled as a branch that waits for each load in turn, and
// loading all eight for every row issues 8x the loads short rows need: the
// bound is the wavefront's longest row, a scalar, and the positions are
// clamped into each row. colIds null: the copies of a row are v[a, a + len)
// (the split operator's slot layout). Every lane of the wavefront must be
// active (the gather loops step by whole wavefronts).
template <typename T>
__device__ inline void gather_rows2(const int* __restrict__ colIds, const T* __restrict__ v, const int (&a)[2],
                                    const int (&len)[2], T (&x)[2][8]) {
  const int wl = wave_max_uniform(max(len[0], len[1]));
  int c[2][8];
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int r = 0; r < 2; ++r)
      c[r][k] = (k < wl) ? (colIds ? colIds[row_copy(a[r], len[r], k)] : row_copy(a[r], len[r], k)) : 0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int r = 0; r < 2; ++r) x[r][k] = (k < wl) ? v[c[r][k]] : T(0);
}

// A workgroup takes row blocks of GB_ROWS rows (gather_block_sums: the copies
// staged in LDS, each row summed from zero in ascending order, the reference's
// order, so results are bit-identical to a thread-per-row gather).
// OP 0: gv = sum; OP 1: gv = r - sum (the residual r - A x of an operator
// whose last step is this gather); OP 2: gv += sum. Each fused form rounds
// once, as the separate axpy it replaces does, so results are bit-identical.
template <typename T, int OP = 0>
__global__ void __launch_bounds__(GB_THREADS) ogs_gather_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                                const int* __restrict__ colIds,
                                                                const T* __restrict__ v, const T* __restrict__ r_vec,
                                                                T* __restrict__ gv) {
  __shared__ T s_v[GB_CAP];
  const long nrb = gather_row_blocks(Nrows);
  for (long rb = blockIdx.x; rb < nrb; rb += gridDim.x) {
    T rr[2] = {T(0), T(0)};
    if (OP != 0) {  // the row's other operand, loaded before the gather's loads complete
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const long row = min(rb * GB_ROWS + threadIdx.x + u * GB_THREADS, (long)Nrows - 1);
        rr[u] = __builtin_nontemporal_load((OP == 1 ? r_vec : gv) + row);
      }
    }
    GatherRows<T> g;
    gather_block_sums(rb, Nrows, rowStarts, colIds, v, s_v, g);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (!g.in[u]) continue;
      const T val = g.val[u];
      T o = val;
      if (OP == 1) o = T(1) * rr[u] + T(-1) * val;
      if (OP == 2) o = T(1) * val + T(1) * rr[u];
      __builtin_nontemporal_store(o, gv + g.row[u]);
    }
  }
}

// Gather fused with a Chebyshev step of the smoother (linalg.hip
// cheb_start_kernel / cheb_step_kernel, ellipticPreconMultiGridLevel.cpp:
// 154-206): the gathered A d (START: A x) of a row is used at once instead of
// being written out and read back. Per row, with a = the gathered value:
//   START: res = invD (r - a); d = invTheta res; RES = res; X += d
//   step:  res = RES - invD a; d = c1 res + c2 d; RES = res; X += d
// The expressions are those of the separate kernels, so results are
// bit-identical.
// The row's operands, loaded before the gather's loads complete (START: invD, r,
// X; step: invD, RES, d, X), with non-temporal accesses as the rows are streamed.
template <typename T, bool START>
struct ChebRow {
  T dv, a, b, x;  // START: a = r; step: a = RES, b = d
};
template <typename T, bool START>
__device__ inline ChebRow<T, START> cheb_row_load(const long row, const T* __restrict__ invDiag,
                                                   const T* __restrict__ rvec, const T* __restrict__ RES,
                                                   const T* __restrict__ d, const T* __restrict__ X) {
  ChebRow<T, START> o;
  o.dv = __builtin_nontemporal_load(invDiag + row);
  o.a = __builtin_nontemporal_load((START ? rvec : RES) + row);
  o.b = START ? T(0) : __builtin_nontemporal_load(d + row);
  o.x = __builtin_nontemporal_load(X + row);
  return o;
}
template <typename T, bool START>
__device__ inline void cheb_row_update(const long row, const T val, const ChebRow<T, START>& q, const T c1,
                                       const T c2, T* __restrict__ RES, T* __restrict__ d, T* __restrict__ X) {
  // no a*b+c contraction, as in cheb_start_kernel / cheb_step_kernel
#pragma clang fp contract(off)
  T res, dn;
  if (START) {
    res = T(1) * q.dv * (T(1) * q.a - T(1) * val);
    dn = c1 * res;
  } else {
    res = T(-1) * q.dv * val + q.a;
    dn = c1 * res + c2 * q.b;
  }
  __builtin_nontemporal_store(res, RES + row);
  __builtin_nontemporal_store(dn, d + row);
  __builtin_nontemporal_store(q.x + dn, X + row);
}

template <typename T, bool START>
__global__ void __launch_bounds__(GB_THREADS) ogs_gather_cheb_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                                     const int* __restrict__ colIds,
                                                                     const T* __restrict__ v,
                                                                     const T* __restrict__ invDiag,
                                                                     const T* __restrict__ rvec, const T c1,
                                                                     const T c2, T* __restrict__ RES,
                                                                     T* __restrict__ d, T* __restrict__ X) {
  __shared__ T s_v[GB_CAP];
  const long nrb = gather_row_blocks(Nrows);
  for (long rb = blockIdx.x; rb < nrb; rb += gridDim.x) {
    ChebRow<T, START> q[2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
      q[u] = cheb_row_load<T, START>(min(rb * GB_ROWS + threadIdx.x + u * GB_THREADS, (long)Nrows - 1), invDiag,
                                     rvec, RES, d, X);
    GatherRows<T> g;
    gather_block_sums(rb, Nrows, rowStarts, colIds, v, s_v, g);
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (g.in[u]) cheb_row_update<T, START>(g.row[u], g.val[u], q[u], c1, c2, RES, d, X);
  }
}

// Gather fused with the p.Ap reduction of a PCG iteration (FP64): gv = gather(v)
// with the arithmetic of ogs_gather_kernel, and sum_row y[row] gv[row]. Each
// block stores one partial in a flat array of up to RNPART*RMAXB entries (the
// ReduceWS partials seen as one buffer), with no fence: the kernel writes gv,
// and a per-block agent-scope fence would write back the dirty L2 of every
// block. gather_pap_finish_kernel then sums the partials in block order and
// writes scalars[PAP] and scalars[ALPHA] = scalars[RDOTZ1] / p.Ap
// (linearSolverPCG.cpp:134-140). Both return at once when scalars[STOP] is set.
constexpr int GATHER_PAP_MAXB = RNPART * RMAXB;

__global__ void __launch_bounds__(GB_THREADS) ogs_gather_pap_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                                    const int* __restrict__ colIds,
                                                                    const double* __restrict__ v,
                                                                    double* __restrict__ gv,
                                                                    const double* __restrict__ y,
                                                                    const double* __restrict__ s,
                                                                    ReduceWS* __restrict__ ws) {
  __shared__ double s_v[GB_CAP];
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const long nrb = gather_row_blocks(Nrows);
  double acc = 0;
  for (long rb = blockIdx.x; rb < nrb; rb += gridDim.x) {
    double yv[2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
      yv[u] = __builtin_nontemporal_load(y + min(rb * GB_ROWS + threadIdx.x + u * GB_THREADS, (long)Nrows - 1));
    GatherRows<double> g;
    gather_block_sums(rb, Nrows, rowStarts, colIds, v, s_v, g);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (!g.in[u]) continue;
      __builtin_nontemporal_store(g.val[u], gv + g.row[u]);
      acc += yv[u] * g.val[u];
    }
  }
  const double b = block_sum(acc, s_w);
  if (threadIdx.x == 0) (&ws->part[0][0])[blockIdx.x] = b;
}

__global__ void __launch_bounds__(256) gather_pap_finish_kernel(const int nb, double* __restrict__ s,
                                                                const ReduceWS* __restrict__ ws) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const double pap = block_sum(strided_sum(&ws->part[0][0], nb), s_w);
  if (threadIdx.x == 0) {
    s[4] = pap;
    s[2] = s[0] / pap;
  }
}

// Gather of the element-boundary rows of the split operator (ax_hex_kernel
// OUT = 1): gv[rows[f]] = the sum of the row's copies in the face E-vector v,
// in ascending local order from zero as in ogs_gather_kernel, so bit-identical
// to the full gather. Rows of element-interior nodes were written by the Ax
// kernel itself. PAP: one partial of sum y[row] gv[row] per block into part.
template <bool PAP>
__global__ void __launch_bounds__(256) ogs_gather_face_kernel(const int Nrows, const int* __restrict__ rows,
                                                              const int* __restrict__ rowStarts,
                                                              const int* __restrict__ colIds,
                                                              const double* __restrict__ v, double* __restrict__ gv,
                                                              const double* __restrict__ y,
                                                              const double* __restrict__ s,
                                                              double* __restrict__ part) {
  __shared__ double s_w[4];
  if (PAP && s[8] != 0.0) return;
  const long S = (long)gridDim.x * blockDim.x;
  double acc = 0;
  for (long w0 = blockIdx.x * (long)blockDim.x + (threadIdx.x & ~63); w0 < Nrows; w0 += 2 * S) {
    const long f0 = w0 + (threadIdx.x & 63);
    int a[2], len[2], row[2];
    double yv[2] = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const long f = f0 + r * S;
      const long fc = min(f, (long)Nrows - 1);  // clamped: the loads of both rows in flight together
      const int rw = rows[fc], s0 = rowStarts[fc], s1 = rowStarts[fc + 1];
      row[r] = (f < Nrows) ? rw : 0;
      a[r] = (f < Nrows) ? s0 : 0;
      len[r] = (f < Nrows) ? s1 - s0 : 0;
    }
    if (PAP) {
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const double yc = y[row[r]];
        yv[r] = (f0 + r * S < Nrows) ? yc : 0.0;
      }
    }
    double x[2][8];
    gather_rows2(colIds, v, a, len, x);  // colIds null: slot layout, copies v[a .. a + len)
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (f0 + r * S >= Nrows) continue;
      double val = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < len[r]) val += x[r][k];
      for (int g = a[r] + 8; g < a[r] + len[r]; ++g) val += v[colIds ? colIds[g] : g];
      gv[row[r]] = val;
      if (PAP) acc += yv[r] * val;
    }
  }
  if (PAP) {
    const double b = block_sum(acc, s_w);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
  }
}

// p.Ap of the split operator: the Ax kernel's interior partials and the face
// gather's partials, summed in a fixed order; then alpha as in
// gather_pap_finish_kernel.
__global__ void __launch_bounds__(1024) split_pap_finish_kernel(const int nb, double* __restrict__ s,
                                                                const double* __restrict__ part) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const double pap = block_sum(strided_sum(part, nb), s_w);
  if (threadIdx.x == 0) {
    s[4] = pap;
    s[2] = s[0] / pap;
  }
}

inline unsigned face_gather_blocks(int Nrows) { return stream_blocks(Nrows, 1024, 4096); }

// The operator's gather fused with the Jacobi-PCG update, for a PCG step whose
// Ax kernel already formed p.Ap and alpha (libp_pcg_ax_hex_pap_f64). Per row, Ap = the sum of the row's copies in
// ascending order from zero (the arithmetic of ogs_gather_kernel), used at
// once: x += alpha p, r -= alpha Ap, z = invDiag.*r, with the expressions of
// pcg_jacobi_update_kernel (linearSolverPCG.cpp:142-160 with
// ellipticPreconJacobi.cpp:42-51). Ap is never stored. Each block stores its
// r.r and r.z partials at part[b] and part[nb + b] (nb = grid size);
// pcg_gather_jacobi_finish_kernel sums them in block order. Both return at
// once while s[8] (STOP) is set. The rows run in row blocks (gather_block_sums),
// one per workgroup (a workgroup loops over row blocks only past
// GATHER_JACOBI_MAXB of them), and the rows' own operands are read and written
// with non-temporal accesses: on the 48^3, N=7 solve (73,429 row blocks) the
// kernel took 501 us against 579 us with 8192 looping workgroups and cached
// accesses (529 us with the non-temporal accesses alone). Non-temporal loads of
// the gathered E-vector were slower (605 us): neighbouring row blocks share its
// cache lines.
constexpr long GATHER_JACOBI_MAXB = 1L << 20;
inline unsigned gather_jacobi_blocks(int Nrows) {
  return (unsigned)std::max<long>(1, std::min<long>(gather_row_blocks(Nrows), GATHER_JACOBI_MAXB));
}

__global__ void __launch_bounds__(GB_THREADS) pcg_gather_jacobi_kernel(const int Nrows,
                                                                       const int* __restrict__ rowStarts,
                                                                       const int* __restrict__ colIds,
                                                                       const double* __restrict__ v,
                                                                       const double* __restrict__ invDiag,
                                                                       const double* __restrict__ p,
                                                                       double* __restrict__ x,
                                                                       double* __restrict__ r,
                                                                       double* __restrict__ z,
                                                                       const double* __restrict__ s,
                                                                       double* __restrict__ part) {
  __shared__ double s_v[GB_CAP];
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const double alpha = s[2];
  const long nrb = gather_row_blocks(Nrows);
  double a0 = 0, a1 = 0;
  for (long rb = blockIdx.x; rb < nrb; rb += gridDim.x) {
    // the rows' own operands, loaded before the gather's loads complete
    double pv[2], xv[2], rv[2], dv[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const long rc = min(rb * GB_ROWS + threadIdx.x + u * GB_THREADS, (long)Nrows - 1);
      pv[u] = __builtin_nontemporal_load(p + rc);
      xv[u] = __builtin_nontemporal_load(x + rc);
      rv[u] = __builtin_nontemporal_load(r + rc);
      dv[u] = __builtin_nontemporal_load(invDiag + rc);
    }
    GatherRows<double> g;
    gather_block_sums(rb, Nrows, rowStarts, colIds, v, s_v, g);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (!g.in[u]) continue;
      const long row = g.row[u];
      const double Ap = g.val[u];
      double xn = xv[u], rn = rv[u];
      xn += alpha * pv[u];
      rn -= alpha * Ap;
      const double zn = 1.0 * dv[u] * rn;
      __builtin_nontemporal_store(xn, x + row);
      __builtin_nontemporal_store(rn, r + row);
      __builtin_nontemporal_store(zn, z + row);
      a0 += rn * rn;
      a1 += rn * zn;
    }
  }
  const double b0 = block_sum(a0, s_w);
  const double b1 = block_sum(a1, s_w);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = b0;
    part[gridDim.x + blockIdx.x] = b1;
  }
}

// First stage of the partial sums when there are many: block b sums entries
// [1024 b, 1024 b + 1024) of both arrays (one load per thread, all in flight)
// into out[b] and out[nb1 + b]. One block summing 73,429 pairs of partials
// that other XCDs wrote took 29 us (about nine rounds of dependent far loads);
// two stages take one round each.
constexpr int GJ_STAGE1 = 1024;
inline int gather_jacobi_stage1_blocks(long nb) { return nb > 2 * GJ_STAGE1 ? (int)((nb + GJ_STAGE1 - 1) / GJ_STAGE1) : 0; }
__global__ void __launch_bounds__(GJ_STAGE1) pcg_gather_jacobi_stage1_kernel(const int nb, const double* __restrict__ s,
                                                                            const double* __restrict__ part,
                                                                            double* __restrict__ out) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const int i = blockIdx.x * GJ_STAGE1 + threadIdx.x;
  const int ic = min(i, nb - 1);
  const double a = part[ic], b = part[nb + ic];
  const double b0 = block_sum(i < nb ? a : 0.0, s_w);
  const double b1 = block_sum(i < nb ? b : 0.0, s_w);
  if (threadIdx.x == 0) {
    out[blockIdx.x] = b0;
    out[gridDim.x + blockIdx.x] = b1;
  }
}

// r.r and r.z from the partials in block order (or from the first stage's
// sums), then ITERS, STOP and the r.z / beta bookkeeping of
// pcg_jacobi_update_finish_kernel.
__global__ void __launch_bounds__(1024) pcg_gather_jacobi_finish_kernel(const int nb, double* __restrict__ s,
                                                                       const double* __restrict__ part) {
  __shared__ double s_w[16];
  if (s[8] != 0.0) return;
  const double rr = block_sum(strided_sum(part, nb), s_w);
  const double rz = block_sum(strided_sum(part + nb, nb), s_w);
  if (threadIdx.x == 0) {
    s[5] = rr;
    s[9] += 1.0;
    if (s[7] > 0.0 && rr <= s[7]) s[8] = 1.0;
    s[1] = s[0];
    s[0] = rz;
    s[3] = rz / s[1];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) ogs_scatter_kernel(const int Nrows, const int* __restrict__ rowStarts,
                                                          const int* __restrict__ colIds, const T* __restrict__ gv,
                                                          T* __restrict__ v) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < Nrows; row += gridDim.x * blockDim.x) {
    const int start = rowStarts[row], end = rowStarts[row + 1];
    const T val = gv[row];
    for (int g = start; g < end; ++g) v[colIds[g]] = val;
  }
}

template <typename T, int OP = 0>
static int gather(int Nrows, const int* rs, const int* ci, const T* v, T* gv, void* stream, const T* r = nullptr) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rs || !ci || !v || !gv || (OP == 1 && !r)) return arg_error("ogs_gather", "null pointer");
  hipLaunchKernelGGL((ogs_gather_kernel<T, OP>), dim3(row_block_grid(Nrows)), dim3(GB_THREADS), 0,
                     as_stream(stream), Nrows, rs, ci, v, r, gv);
  return launch_status("ogs_gather");
}

template <typename T>
static int scatter(int Nrows, const int* rs, const int* ci, const T* gv, T* v, void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rs || !ci || !v || !gv) return arg_error("ogs_scatter", "null pointer");
  hipLaunchKernelGGL(ogs_scatter_kernel<T>, dim3(stream_blocks(Nrows, 256)), dim3(256), 0, as_stream(stream),
                     Nrows, rs, ci, gv, v);
  return launch_status("ogs_scatter");
}

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_ogs_gather_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v, double* gv,
                                   void* stream) {
  return gather<double>(Nrows, rowStarts, colIds, v, gv, stream);
}
extern "C" int libp_ogs_gather_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v, float* gv,
                                   void* stream) {
  return gather<float>(Nrows, rowStarts, colIds, v, gv, stream);
}
extern "C" int libp_ogs_scatter_f64(int Nrows, const int* rowStarts, const int* colIds, const double* gv, double* v,
                                    void* stream) {
  return scatter<double>(Nrows, rowStarts, colIds, gv, v, stream);
}
extern "C" int libp_ogs_scatter_i32(int Nrows, const int* rowStarts, const int* colIds, const int* gv, int* v,
                                    void* stream) {
  return scatter<int>(Nrows, rowStarts, colIds, gv, v, stream);
}

extern "C" int libp_pcg_gather_pap_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v,
                                       double* gv, const double* p, double* scalars, void* ws, void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rowStarts || !colIds || !v || !gv || !p || !scalars || !ws) return arg_error("pcg_gather_pap", "null pointer");
  const unsigned blocks = (unsigned)std::min<long>(gather_row_blocks(Nrows), GATHER_PAP_MAXB);
  hipLaunchKernelGGL(ogs_gather_pap_kernel, dim3(blocks), dim3(GB_THREADS), 0, as_stream(stream), Nrows, rowStarts,
                     colIds, v, gv, p, scalars, (ReduceWS*)ws);
  hipLaunchKernelGGL(gather_pap_finish_kernel, dim3(1), dim3(256), 0, as_stream(stream), (int)blocks, scalars,
                     (const ReduceWS*)ws);
  return launch_status("pcg_gather_pap");
}

extern "C" int libp_ogs_gather_residual_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v,
                                            const double* r, double* gv, void* stream) {
  return gather<double, 1>(Nrows, rowStarts, colIds, v, gv, stream, r);
}
extern "C" int libp_ogs_gather_residual_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v,
                                            const float* r, float* gv, void* stream) {
  return gather<float, 1>(Nrows, rowStarts, colIds, v, gv, stream, r);
}
extern "C" int libp_ogs_gather_add_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v, double* gv,
                                       void* stream) {
  return gather<double, 2>(Nrows, rowStarts, colIds, v, gv, stream);
}
extern "C" int libp_ogs_gather_add_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v, float* gv,
                                       void* stream) {
  return gather<float, 2>(Nrows, rowStarts, colIds, v, gv, stream);
}

template <typename T, bool START>
static int gather_cheb(int Nrows, const int* rs, const int* ci, const T* v, const T* invDiag, const T* r, T c1, T c2,
                       T* RES, T* d, T* X, void* stream) {
  if (Nrows <= 0) return LIBP_OK;
  if (!rs || !ci || !v || !invDiag || !RES || !d || !X || (START && !r))
    return arg_error("ogs_gather_cheb", "null pointer");
  hipLaunchKernelGGL((ogs_gather_cheb_kernel<T, START>), dim3(row_block_grid(Nrows)), dim3(GB_THREADS), 0,
                     as_stream(stream), Nrows, rs, ci, v, invDiag, r, c1, c2, RES, d, X);
  return launch_status("ogs_gather_cheb");
}
extern "C" int libp_ogs_gather_cheb_start_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v,
                                              const double* invDiag, const double* r, double invTheta, double* RES,
                                              double* d, double* X, void* stream) {
  return gather_cheb<double, true>(Nrows, rowStarts, colIds, v, invDiag, r, invTheta, 0.0, RES, d, X, stream);
}
extern "C" int libp_ogs_gather_cheb_start_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v,
                                              const float* invDiag, const float* r, float invTheta, float* RES,
                                              float* d, float* X, void* stream) {
  return gather_cheb<float, true>(Nrows, rowStarts, colIds, v, invDiag, r, invTheta, 0.f, RES, d, X, stream);
}
extern "C" int libp_ogs_gather_cheb_step_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v,
                                             const double* invDiag, double c1, double c2, double* RES, double* d,
                                             double* X, void* stream) {
  return gather_cheb<double, false>(Nrows, rowStarts, colIds, v, invDiag, nullptr, c1, c2, RES, d, X, stream);
}
extern "C" int libp_ogs_gather_cheb_step_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v,
                                             const float* invDiag, float c1, float c2, float* RES, float* d, float* X,
                                             void* stream) {
  return gather_cheb<float, false>(Nrows, rowStarts, colIds, v, invDiag, nullptr, c1, c2, RES, d, X, stream);
}

extern "C" int libp_ogs_gather_face_blocks(int NfaceRows) { return NfaceRows > 0 ? (int)face_gather_blocks(NfaceRows) : 0; }

extern "C" int libp_ogs_gather_face_f64(int NfaceRows, const int* faceRows, const int* faceRowStarts,
                                        const int* faceCols, const double* AqFace, double* Aq, void* stream) {
  if (NfaceRows <= 0) return LIBP_OK;
  if (!faceRows || !faceRowStarts || !AqFace || !Aq) return arg_error("ogs_gather_face", "null pointer");
  hipLaunchKernelGGL(ogs_gather_face_kernel<false>, dim3(face_gather_blocks(NfaceRows)), dim3(256), 0,
                     as_stream(stream), NfaceRows, faceRows, faceRowStarts, faceCols, AqFace, Aq, nullptr, nullptr,
                     nullptr);
  return launch_status("ogs_gather_face");
}

extern "C" int libp_pcg_gather_face_pap_f64(int NfaceRows, const int* faceRows, const int* faceRowStarts,
                                            const int* faceCols, const double* AqFace, double* Ap, const double* p,
                                            int nAxPartials, double* partials, double* scalars, void* stream) {
  if (!faceRows || !faceRowStarts || !AqFace || !Ap || !p || !partials || !scalars || nAxPartials < 0 ||
      NfaceRows < 0)
    return arg_error("pcg_gather_face_pap", "null pointer or negative count");
  const unsigned blocks = NfaceRows > 0 ? face_gather_blocks(NfaceRows) : 0;
  if (blocks > 0)
    hipLaunchKernelGGL(ogs_gather_face_kernel<true>, dim3(blocks), dim3(256), 0, as_stream(stream), NfaceRows,
                       faceRows, faceRowStarts, faceCols, AqFace, Ap, p, scalars, partials + nAxPartials);
  hipLaunchKernelGGL(split_pap_finish_kernel, dim3(1), dim3(1024), 0, as_stream(stream), nAxPartials + (int)blocks,
                     scalars, partials);
  return launch_status("pcg_gather_face_pap");
}

extern "C" long libp_pcg_gather_jacobi_partials(int Nrows) {
  if (Nrows < 0) return -1;
  const long nb = gather_jacobi_blocks(Nrows);
  return 2L * nb + 2L * gather_jacobi_stage1_blocks(nb);  // the blocks' partials, then the first stage's sums
}

extern "C" int libp_pcg_gather_jacobi_update_f64(int Nrows, const int* rowStarts, const int* colIds,
                                                 const double* AqL, const double* invDiag, const double* p, double* x,
                                                 double* r, double* z, double* scalars, double* partials,
                                                 void* stream) {
  if (Nrows < 0) return arg_error("pcg_gather_jacobi_update", "negative row count");
  if (!rowStarts || !colIds || !AqL || !invDiag || !p || !x || !r || !z || !scalars || !partials)
    return arg_error("pcg_gather_jacobi_update", "null pointer");
  const unsigned nb = gather_jacobi_blocks(Nrows);
  hipLaunchKernelGGL(pcg_gather_jacobi_kernel, dim3(nb), dim3(GB_THREADS), 0, as_stream(stream), Nrows, rowStarts,
                     colIds, AqL, invDiag, p, x, r, z, (const double*)scalars, partials);
  const int nb1 = gather_jacobi_stage1_blocks(nb);
  if (nb1 > 0) {
    double* sums = partials + 2L * nb;
    hipLaunchKernelGGL(pcg_gather_jacobi_stage1_kernel, dim3(nb1), dim3(GJ_STAGE1), 0, as_stream(stream), (int)nb,
                       (const double*)scalars, (const double*)partials, sums);
    hipLaunchKernelGGL(pcg_gather_jacobi_finish_kernel, dim3(1), dim3(1024), 0, as_stream(stream), nb1, scalars,
                       (const double*)sums);
  } else {
    hipLaunchKernelGGL(pcg_gather_jacobi_finish_kernel, dim3(1), dim3(1024), 0, as_stream(stream), (int)nb, scalars,
                       (const double*)partials);
  }
  return launch_status("pcg_gather_jacobi_update");
}
