// BF16 GEMM tiles on v_mfma_f32_32x32x16_bf16 (FP32 accumulate) for the
// learned per-element smoother: forward layers, backward data and weight
// gradients, split-K partial sums, column sums and the Adam step. No
// reference counterpart (the reference has no learned components).
//
// Operand layout of v_mfma_f32_32x32x16_bf16 (C[32x32] += A[32x16] B[16x32],
// wave64): lane l supplies eight A values of row l%32 and eight B values of
// column l%32, the K slice 8*(l/32)..8*(l/32)+7; it holds
// c[r] = C[8*(r/4) + 4*(l/32) + r%4][l%32]. A and B take the same K slice per
// lane, so the product does not depend on how the hardware orders K inside it.
//
// Tiles: 128x128 per workgroup of four wavefronts (2x2, 64x64 each = 2x2
// MFMA blocks), BK = 32, two LDS buffers with register prefetch of the next
// K tile. Operands reach LDS as [row][k] (row stride 40 elements) whatever
// their layout in HBM: a K-contiguous operand is copied with 16-byte loads and
// stores, a row-contiguous one (the transposed operands of the weight
// gradients) with 16-byte loads and 2-byte LDS stores.
#include <cmath>

#include "reduce.hpp"

namespace libp_amd {

typedef unsigned short u16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even; NaN stays NaN
__device__ inline u16 f2bf(float f) {
  unsigned u = __builtin_bit_cast(unsigned, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ inline float bf2f(u16 h) { return __builtin_bit_cast(float, (unsigned)h << 16); }

constexpr int GBM = 128, GBN = 128, GBK = 32, GLD = GBK + 8;
constexpr int GTHREADS = 256;

enum { EPI_F32_BIAS = 0, EPI_BF16_BIAS_RELU = 1, EPI_BF16_MASK = 2, EPI_F32_PARTIAL = 3 };

struct GemmArgs {
  int M, N, K;          // C is M x N, reduction over K
  const u16* A;         // TA=false: A[m*lda + k]; TA=true: A[k*lda + m]
  long lda;
  const u16* B;         // TB=false: B[n*ldb + k]; TB=true: B[k*ldb + n]
  long ldb;
  void* C;              // [M][ldc]; split s of a partial output at C + s*M*ldc (FP32)
  long ldc;
  const float* bias;    // [N] or null
  const u16* mask;      // [M][ldm] or null: output zeroed where mask <= 0
  long ldm;
  int kChunk;           // K range of one split (multiple of GBK)
  const float* rowScale = nullptr;  // EPI_F32_BIAS: C[m][n] = rowScale[m] * (A B + bias)[m][n] when set
};

// One operand tile (rows x GBK) in registers: two 16-byte chunks per thread.
struct TileRegs {
  u32x4 v[2];
};

template <bool T>
__device__ inline void load_tile(TileRegs& R, const u16* __restrict__ X, long ld, int row0, int nrows, int k0, int kEnd,
                                 bool vec) {
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = t + GTHREADS * i;
    u16 tmp[8];
    if (!T) {
      const int row = c >> 2, kc = (c & 3) * 8;
      const int gr = row0 + row, gk = k0 + kc;
      if (vec && gr < nrows && gk + 8 <= kEnd) {
        R.v[i] = *reinterpret_cast<const u32x4*>(X + (long)gr * ld + gk);
        continue;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) tmp[e] = (gr < nrows && gk + e < kEnd) ? X[(long)gr * ld + gk + e] : (u16)0;
    } else {
      const int kr = c >> 4, rc = (c & 15) * 8;
      const int gk = k0 + kr, gr = row0 + rc;
      if (vec && gk < kEnd && gr + 8 <= nrows) {
        R.v[i] = *reinterpret_cast<const u32x4*>(X + (long)gk * ld + gr);
        continue;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) tmp[e] = (gk < kEnd && gr + e < nrows) ? X[(long)gk * ld + gr + e] : (u16)0;
    }
    R.v[i] = u32x4{(unsigned)tmp[0] | ((unsigned)tmp[1] << 16), (unsigned)tmp[2] | ((unsigned)tmp[3] << 16),
                   (unsigned)tmp[4] | ((unsigned)tmp[5] << 16), (unsigned)tmp[6] | ((unsigned)tmp[7] << 16)};
  }
}

template <bool T>
__device__ inline void store_tile(const TileRegs& R, u16* __restrict__ s) {
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = t + GTHREADS * i;
    if (!T) {
      const int row = c >> 2, kc = (c & 3) * 8;
      *reinterpret_cast<u32x4*>(s + row * GLD + kc) = R.v[i];
    } else {
      const int kr = c >> 4, rc = (c & 15) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const unsigned w = R.v[i][e >> 1];
        s[(rc + e) * GLD + kr] = (u16)((e & 1) ? (w >> 16) : (w & 0xffffu));
      }
    }
  }
}

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(GTHREADS, 4) gemm_bf16_kernel(const GemmArgs g) {
  __shared__ __attribute__((aligned(16))) u16 sA[2][GBM * GLD];
  __shared__ __attribute__((aligned(16))) u16 sB[2][GBN * GLD];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int wm = w >> 1, wn = w & 1;
  // Workgroups are dispatched round-robin over the eight XCDs, each with its
  // own L2. When the M tiles come in multiples of eight, the tiles are renamed
  // so that the gridDim.x N tiles of one M tile run on one XCD, one after
  // another: the A rows (the large operand) then reach that L2 once instead of
  // once per N tile. The tiles and their arithmetic are unchanged.
  int mt = blockIdx.y, nt = blockIdx.x;
  if ((gridDim.y & 7) == 0) {
    const int b = blockIdx.x + gridDim.x * blockIdx.y;  // dispatch order within a K split
    const int k = b >> 3;
    nt = k % gridDim.x;
    mt = (k / gridDim.x) * 8 + (b & 7);
  }
  const int m0 = mt * GBM, n0 = nt * GBN;
  const int kBeg = blockIdx.z * g.kChunk;
  const int kEnd = min(g.K, kBeg + g.kChunk);
  const bool vecA = (g.lda % 8) == 0 && ((uintptr_t)g.A & 15) == 0;
  const bool vecB = (g.ldb % 8) == 0 && ((uintptr_t)g.B & 15) == 0;

  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  const int nT = (kEnd > kBeg) ? (kEnd - kBeg + GBK - 1) / GBK : 0;
  TileRegs ra, rb;
  if (nT > 0) {
    load_tile<TA>(ra, g.A, g.lda, m0, g.M, kBeg, kEnd, vecA);
    load_tile<TB>(rb, g.B, g.ldb, n0, g.N, kBeg, kEnd, vecB);
    store_tile<TA>(ra, sA[0]);
    store_tile<TB>(rb, sB[0]);
  }
  __syncthreads();
  const int lr = l & 31, lk = 8 * (l >> 5);
  for (int it = 0; it < nT; ++it) {
    const int buf = it & 1;
    if (it + 1 < nT) {
      const int k0 = kBeg + (it + 1) * GBK;
      load_tile<TA>(ra, g.A, g.lda, m0, g.M, k0, kEnd, vecA);
      load_tile<TB>(rb, g.B, g.ldb, n0, g.N, k0, kEnd, vecB);
    }
    const u16* a_s = sA[buf];
    const u16* b_s = sB[buf];
#pragma unroll
    for (int ks = 0; ks < GBK; ks += 16) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[i] = *reinterpret_cast<const bf16x8*>(a_s + (wm * 64 + i * 32 + lr) * GLD + ks + lk);
#pragma unroll
      for (int j = 0; j < 2; ++j)
        b[j] = *reinterpret_cast<const bf16x8*>(b_s + (wn * 64 + j * 32 + lr) * GLD + ks + lk);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (it + 1 < nT) {
      store_tile<TA>(ra, sA[buf ^ 1]);
      store_tile<TB>(rb, sB[buf ^ 1]);
    }
    __syncthreads();
  }

  // epilogue
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn * 64 + j * 32 + lr;
      if (col >= g.N) continue;
      const float bcol = (g.bias && (EPI == EPI_F32_BIAS || EPI == EPI_BF16_BIAS_RELU)) ? g.bias[col] : 0.f;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
        if (row >= g.M) continue;
        const float v = acc[i][j][r];
        if (EPI == EPI_F32_BIAS) {
          const float o = v + bcol;
          reinterpret_cast<float*>(g.C)[(long)row * g.ldc + col] = g.rowScale ? g.rowScale[row] * o : o;
        } else if (EPI == EPI_BF16_BIAS_RELU) {
          reinterpret_cast<u16*>(g.C)[(long)row * g.ldc + col] = f2bf(fmaxf(v + bcol, 0.f));
        } else if (EPI == EPI_BF16_MASK) {
          const bool keep = !g.mask || bf2f(g.mask[(long)row * g.ldm + col]) > 0.f;
          reinterpret_cast<u16*>(g.C)[(long)row * g.ldc + col] = f2bf(keep ? v : 0.f);
        } else {
          reinterpret_cast<float*>(g.C)[(long)blockIdx.z * g.M * g.ldc + (long)row * g.ldc + col] = v;
        }
      }
    }
}

template <bool TA, bool TB, int EPI>
static int gemm(const GemmArgs& g, int splits, void* stream, const char* where) {
  if (g.M <= 0 || g.N <= 0) return LIBP_OK;
  dim3 grid((g.N + GBN - 1) / GBN, (g.M + GBM - 1) / GBM, splits);
  hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, EPI>), grid, dim3(GTHREADS), 0, as_stream(stream), g);
  return launch_status(where);
}

// out[i] (+)= sum_s P[s*n + i], s in order (deterministic)
__global__ void __launch_bounds__(256) sum_splits_kernel(const long n, const int S, const float* __restrict__ P,
                                                         float* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float a = 0.f;
    for (int s = 0; s < S; ++s) a += P[(long)s * n + i];
    out[i] = a;
  }
}

// partial column sums of a BF16 [M][ld] matrix: P[chunk][col]
__global__ void __launch_bounds__(256) colsum_bf16_kernel(const int M, const int N, const u16* __restrict__ X,
                                                          const long ld, const int rowsPerChunk,
                                                          float* __restrict__ P) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= N) return;
  const int r0 = blockIdx.y * rowsPerChunk, r1 = min(M, r0 + rowsPerChunk);
  float a = 0.f;
  for (int r = r0; r < r1; ++r) a += bf2f(X[(long)r * ld + col]);
  P[(long)blockIdx.y * N + col] = a;
}

__global__ void __launch_bounds__(256) f32_to_bf16_kernel(const long n, const float* __restrict__ x,
                                                          u16* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// Adam (Kingma & Ba 2015) on FP32 masters with a BF16 copy of the result.
// c1 = 1/(1 - beta1^t), c2 = 1/(1 - beta2^t).
__global__ void __launch_bounds__(256) adam_kernel(const int n, float* __restrict__ p, u16* __restrict__ pb,
                                                   const float* __restrict__ grad, float* __restrict__ m,
                                                   float* __restrict__ v, const float lr, const float b1,
                                                   const float b2, const float eps, const float c1, const float c2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = f2bf(pi);
  }
}

// split-K sizing of the weight gradient (reduction over M rows)
static void weight_splits(int M, int& S, int& kChunk) {
  S = (M + 1023) / 1024;
  if (S < 1) S = 1;
  if (S > 64) S = 64;
  kChunk = ((M + S - 1) / S + GBK - 1) / GBK * GBK;
  S = (M + kChunk - 1) / kChunk;
  if (S < 1) S = 1;
}
static int colsum_chunks(int M) { return M < 256 ? 1 : (M + 255) / 256 > 128 ? 128 : (M + 255) / 256; }

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_mlp_linear_fwd(int M, int K, int Nout, const unsigned short* X, long ldx, const unsigned short* Wt,
                                   long ldw, const float* b, int relu, void* Y, long ldy, void* stream) {
  if (M < 0 || K <= 0 || Nout <= 0) return arg_error("mlp_linear_fwd", "bad size");
  if (!X || !Wt || !Y) return arg_error("mlp_linear_fwd", "null pointer");
  if (ldx < K || ldw < K || ldy < Nout) return arg_error("mlp_linear_fwd", "leading dimension too small");
  GemmArgs g{M, Nout, K, X, ldx, Wt, ldw, Y, ldy, b, nullptr, 0, ((K + GBK - 1) / GBK) * GBK};
  if (relu) return gemm<false, false, EPI_BF16_BIAS_RELU>(g, 1, stream, "mlp_linear_fwd");
  return gemm<false, false, EPI_F32_BIAS>(g, 1, stream, "mlp_linear_fwd");
}

extern "C" int libp_mlp_linear_fwd_scaled(int M, int K, int Nout, const unsigned short* X, long ldx,
                                          const unsigned short* Wt, long ldw, const float* b, const float* rowScale,
                                          float* Y, long ldy, void* stream) {
  if (M < 0 || K <= 0 || Nout <= 0) return arg_error("mlp_linear_fwd_scaled", "bad size");
  if (!X || !Wt || !Y || !rowScale) return arg_error("mlp_linear_fwd_scaled", "null pointer");
  if (ldx < K || ldw < K || ldy < Nout) return arg_error("mlp_linear_fwd_scaled", "leading dimension too small");
  GemmArgs g{M, Nout, K, X, ldx, Wt, ldw, Y, ldy, b, nullptr, 0, ((K + GBK - 1) / GBK) * GBK, rowScale};
  return gemm<false, false, EPI_F32_BIAS>(g, 1, stream, "mlp_linear_fwd_scaled");
}

extern "C" int libp_mlp_linear_bwd_data(int M, int K, int Nout, const unsigned short* dY, long lddy,
                                        const unsigned short* Wt, long ldw, const unsigned short* H, long ldh,
                                        unsigned short* dX, long lddx, void* stream) {
  if (M < 0 || K <= 0 || Nout <= 0) return arg_error("mlp_linear_bwd_data", "bad size");
  if (!dY || !Wt || !dX) return arg_error("mlp_linear_bwd_data", "null pointer");
  if (lddy < Nout || ldw < K || lddx < K || (H && ldh < K)) return arg_error("mlp_linear_bwd_data", "leading dimension");
  // dX[m][k] = sum_n dY[m][n] Wt[n][k]: A = dY (K-contiguous over n), B(k, n) = Wt[n*ldw + k] (transposed)
  GemmArgs g{M, K, Nout, dY, lddy, Wt, ldw, dX, lddx, nullptr, H, ldh, ((Nout + GBK - 1) / GBK) * GBK};
  return gemm<false, true, EPI_BF16_MASK>(g, 1, stream, "mlp_linear_bwd_data");
}

extern "C" long libp_mlp_ws_bytes(int M, int K, int Nout) {
  int S, kc;
  weight_splits(M, S, kc);
  return (long)S * Nout * K * 4 + (long)colsum_chunks(M) * Nout * 4 + 256;
}

extern "C" int libp_mlp_linear_bwd_weight(int M, int K, int Nout, const unsigned short* dY, long lddy,
                                          const unsigned short* X, long ldx, float* dWt, float* db, void* ws,
                                          void* stream) {
  if (M <= 0 || K <= 0 || Nout <= 0) return arg_error("mlp_linear_bwd_weight", "bad size");
  if (!dY || !X || !dWt || !ws) return arg_error("mlp_linear_bwd_weight", "null pointer");
  if (lddy < Nout || ldx < K) return arg_error("mlp_linear_bwd_weight", "leading dimension");
  int S, kc;
  weight_splits(M, S, kc);
  float* P = reinterpret_cast<float*>(ws);
  // dWt[n][k] = sum_m dY[m][n] X[m][k]: A(n, m) = dY[m*lddy + n], B(k, m) = X[m*ldx + k] (both transposed)
  GemmArgs g{Nout, K, M, dY, lddy, X, ldx, P, K, nullptr, nullptr, 0, kc};
  int st = gemm<true, true, EPI_F32_PARTIAL>(g, S, stream, "mlp_linear_bwd_weight");
  if (st) return st;
  const long n = (long)Nout * K;
  hipLaunchKernelGGL(sum_splits_kernel, dim3(stream_blocks(n, 256, 4096)), dim3(256), 0, as_stream(stream), n, S,
                     P, dWt);
  if ((st = launch_status("mlp_linear_bwd_weight"))) return st;
  if (db) {
    const int C = colsum_chunks(M);
    float* Pc = P + (long)S * Nout * K;
    hipLaunchKernelGGL(colsum_bf16_kernel, dim3((Nout + 255) / 256, C), dim3(256), 0, as_stream(stream), M, Nout, dY,
                       lddy, (M + C - 1) / C, Pc);
    hipLaunchKernelGGL(sum_splits_kernel, dim3(stream_blocks(Nout, 256, 64)), dim3(256), 0, as_stream(stream),
                       (long)Nout, C, Pc, db);
    st = launch_status("mlp_linear_bwd_weight");
  }
  return st;
}

extern "C" int libp_adam_step(int n, float* master, unsigned short* wbf16, const float* grad, float* m, float* v,
                              float lr, float beta1, float beta2, float eps, int step, void* stream) {
  if (n <= 0) return LIBP_OK;
  if (!master || !grad || !m || !v) return arg_error("adam_step", "null pointer");
  if (step < 1) return arg_error("adam_step", "step counts from 1");
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, step)));
  hipLaunchKernelGGL(adam_kernel, dim3(stream_blocks(n, 256, 4096)), dim3(256), 0, as_stream(stream), n, master, wbf16,
                     grad, m, v, lr, beta1, beta2, eps, c1, c2);
  return launch_status("adam_step");
}

extern "C" int libp_f32_to_bf16(long n, const float* x, unsigned short* y, void* stream) {
  if (n <= 0) return LIBP_OK;
  if (!x || !y) return arg_error("f32_to_bf16", "null pointer");
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(stream_blocks(n, 256, 4096)), dim3(256), 0, as_stream(stream), n, x, y);
  return launch_status("f32_to_bf16");
}
