// Setup-time kernels: physical nodes and geometric factors, operator
// diagonal, forcing and boundary data. They run once per level, but on the
// 48^3 x N=7 box they replace ~3.6 GB of host-side array generation and its
// PCIe upload, so they are done where the data is consumed.
#include "kcommon.hpp"

namespace libp_amd {

constexpr double kPI = 3.14159265358979323846;

// meshPhysicalNodesHex3D.cpp:31-112 (+ curved map x += amp sin(2 pi y) sin(2 pi z)).
__global__ void __launch_bounds__(256) nodes_kernel(const int Nq, const long NL, const double* __restrict__ EX,
                                                    const double* __restrict__ EY, const double* __restrict__ EZ,
                                                    const double* __restrict__ r1, const double curveAmp,
                                                    double* __restrict__ x, double* __restrict__ y,
                                                    double* __restrict__ z) {
  const int Np = Nq * Nq * Nq;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < NL; id += (long)gridDim.x * blockDim.x) {
    const long e = id / Np;
    const int n = (int)(id - e * Np);
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    const double rn = r1[i], sn = r1[j], tn = r1[k];
    const double c[8] = {0.125 * (1 - rn) * (1 - sn) * (1 - tn), 0.125 * (1 + rn) * (1 - sn) * (1 - tn),
                         0.125 * (1 + rn) * (1 + sn) * (1 - tn), 0.125 * (1 - rn) * (1 + sn) * (1 - tn),
                         0.125 * (1 - rn) * (1 - sn) * (1 + tn), 0.125 * (1 + rn) * (1 - sn) * (1 + tn),
                         0.125 * (1 + rn) * (1 + sn) * (1 + tn), 0.125 * (1 - rn) * (1 + sn) * (1 + tn)};
    double X = 0, Y = 0, Z = 0;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      X += c[v] * EX[e * 8 + v];
      Y += c[v] * EY[e * 8 + v];
      Z += c[v] * EZ[e * 8 + v];
    }
    if (curveAmp != 0.0) X += curveAmp * sin(2 * kPI * Y) * sin(2 * kPI * Z);
    x[id] = X;
    y[id] = Y;
    z[id] = Z;
  }
}

// meshGeometricFactorsHex3D.cpp:84-171 (ggeo and wJ only).
__global__ void __launch_bounds__(256) geofacs_kernel(const int Nq, const long NL, const double* __restrict__ w1,
                                                      const double* __restrict__ D, const double* __restrict__ x,
                                                      const double* __restrict__ y, const double* __restrict__ z,
                                                      double* __restrict__ ggeo, double* __restrict__ wJ,
                                                      int* __restrict__ badElement) {
  const int Np = Nq * Nq * Nq;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < NL; id += (long)gridDim.x * blockDim.x) {
    const long e = id / Np;
    const int n = (int)(id - e * Np);
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    double xr = 0, xs = 0, xt = 0, yr = 0, ys = 0, yt = 0, zr = 0, zs = 0, zt = 0;
    const long eb = e * Np;
    for (int m = 0; m < Nq; ++m) {
      const long idr = eb + k * Nq * Nq + j * Nq + m;
      const long ids = eb + k * Nq * Nq + m * Nq + i;
      const long idt = eb + m * Nq * Nq + j * Nq + i;
      const double Di = D[i * Nq + m], Dj = D[j * Nq + m], Dk = D[k * Nq + m];
      xr += Di * x[idr]; xs += Dj * x[ids]; xt += Dk * x[idt];
      yr += Di * y[idr]; ys += Dj * y[ids]; yt += Dk * y[idt];
      zr += Di * z[idr]; zs += Dj * z[ids]; zt += Dk * z[idt];
    }
    const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
    if (J < 1e-12) atomicMax(badElement, (int)e + 1);
    const double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
    const double sx = -(yr * zt - zr * yt) / J, sy = (xr * zt - zr * xt) / J, sz = -(xr * yt - yr * xt) / J;
    const double tx = (yr * zs - zr * ys) / J, ty = -(xr * zs - zr * xs) / J, tz = (xr * ys - yr * xs) / J;
    const double JW = J * w1[i] * w1[j] * w1[k];
    double* g = ggeo + e * 6 * Np + n;
    g[0 * Np] = JW * (rx * rx + ry * ry + rz * rz);
    g[1 * Np] = JW * (rx * sx + ry * sy + rz * sz);
    g[2 * Np] = JW * (rx * tx + ry * ty + rz * tz);
    g[3 * Np] = JW * (sx * sx + sy * sy + sz * sz);
    g[4 * Np] = JW * (sx * tx + sy * ty + sz * tz);
    g[5 * Np] = JW * (tx * tx + ty * ty + tz * tz);
    wJ[id] = JW;
  }
}

// BuildOperatorDiagonalContinuousHex3D (ellipticBuildOperatorDiagonal.cpp:998-1057).
__global__ void __launch_bounds__(256) diag_kernel(const int Nq, const long NL, const double* __restrict__ ggeo,
                                                   const double* __restrict__ wJ, const double* __restrict__ D,
                                                   const double lambda, const int* __restrict__ mapB,
                                                   const int allNeumann, const double shift,
                                                   double* __restrict__ diagL) {
  const int Np = Nq * Nq * Nq;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < NL; id += (long)gridDim.x * blockDim.x) {
    const long e = id / Np;
    const int n = (int)(id - e * Np);
    const int nx = n % Nq, ny = (n / Nq) % Nq, nz = n / (Nq * Nq);
    double a;
    if (mapB[id] != 1) {
      const double* g = ggeo + e * 6 * Np;
      a = 0;
      a += 2 * g[n + 1 * Np] * D[nx + nx * Nq] * D[ny + ny * Nq];
      a += 2 * g[n + 2 * Np] * D[nx + nx * Nq] * D[nz + nz * Nq];
      a += 2 * g[n + 4 * Np] * D[ny + ny * Nq] * D[nz + nz * Nq];
      for (int k = 0; k < Nq; k++) a += g[k + ny * Nq + nz * Nq * Nq + 0 * Np] * D[nx + k * Nq] * D[nx + k * Nq];
      for (int k = 0; k < Nq; k++) a += g[nx + k * Nq + nz * Nq * Nq + 3 * Np] * D[ny + k * Nq] * D[ny + k * Nq];
      for (int k = 0; k < Nq; k++) a += g[nx + ny * Nq + k * Nq * Nq + 5 * Np] * D[nz + k * Nq] * D[nz + k * Nq];
      a += wJ[id] * lambda;
      if (allNeumann) a += shift;
    } else {
      a = 1;
    }
    diagL[id] = a;
  }
}

struct Data {
  int kind;
  const double* c;  // 64 x {kx,ky,kz,amp,phase}
  __device__ double forcing(double x, double y, double z, double lambda) const {
    if (kind == 0) return (3 * kPI * kPI + lambda) * sin(kPI * x) * sin(kPI * y) * sin(kPI * z);
    double f = 0;
    for (int m = 0; m < 64; ++m) {
      const double* cm = c + 5 * m;
      f += cm[3] * sin(kPI * (cm[0] * x + cm[1] * y + cm[2] * z) + cm[4]);
    }
    return f;
  }
  __device__ double dirichlet(double x, double y, double z) const {
    if (kind == 0) return sin(kPI * x) * sin(kPI * y) * sin(kPI * z);
    return 0.0;
  }
};

__global__ void __launch_bounds__(256) rhs_kernel(const long NL, const Data data, const double* __restrict__ x,
                                                  const double* __restrict__ y, const double* __restrict__ z,
                                                  const double* __restrict__ wJ, const double lambda,
                                                  const int* __restrict__ mapB, double* __restrict__ rhsL,
                                                  double* __restrict__ ubL) {
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < NL; id += (long)gridDim.x * blockDim.x) {
    const double X = x[id], Y = y[id], Z = z[id];
    rhsL[id] = wJ[id] * data.forcing(X, Y, Z, lambda);
    if (ubL) ubL[id] = (mapB[id] == 1) ? data.dirichlet(X, Y, Z) : 0.0;
  }
}

__global__ void __launch_bounds__(256) addbc_kernel(const long NL, const Data data, const double* __restrict__ x,
                                                    const double* __restrict__ y, const double* __restrict__ z,
                                                    const int* __restrict__ mapB, double* __restrict__ xL) {
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < NL; id += (long)gridDim.x * blockDim.x)
    if (mapB[id] == 1) xL[id] = data.dirichlet(x[id], y[id], z[id]);
}

}  // namespace libp_amd

using namespace libp_amd;

extern "C" int libp_geofacs_hex_f64(int N, int Nelements, const double* EX, const double* EY, const double* EZ,
                                    const double* r1, const double* w1, const double* D, double curveAmp, double* x,
                                    double* y, double* z, double* ggeo, double* wJ, int* badElement, void* stream) {
  if (N < 1 || N > 15) return arg_error("geofacs_hex", "degree out of range");
  if (Nelements <= 0) return LIBP_OK;
  const int Nq = N + 1;
  const long NL = (long)Nelements * Nq * Nq * Nq;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(nodes_kernel, dim3(stream_blocks(NL, 256)), dim3(256), 0, st, Nq, NL, EX, EY, EZ, r1, curveAmp,
                     x, y, z);
  hipLaunchKernelGGL(geofacs_kernel, dim3(stream_blocks(NL, 256)), dim3(256), 0, st, Nq, NL, w1, D, x, y, z, ggeo,
                     wJ, badElement);
  return launch_status("geofacs_hex");
}

extern "C" int libp_diag_hex_local_f64(int N, int Nelements, const double* ggeo, const double* wJ, const double* D,
                                       double lambda, const int* mapB, int allNeumann, double neumannShift,
                                       double* diagL, void* stream) {
  if (N < 1 || N > 15) return arg_error("diag_hex", "degree out of range");
  const int Nq = N + 1;
  const long NL = (long)Nelements * Nq * Nq * Nq;
  if (NL <= 0) return LIBP_OK;
  hipLaunchKernelGGL(diag_kernel, dim3(stream_blocks(NL, 256)), dim3(256), 0, as_stream(stream), Nq, NL, ggeo, wJ, D,
                     lambda, mapB, allNeumann, neumannShift, diagL);
  return launch_status("diag_hex");
}

extern "C" int libp_rhs_hex_f64(int Nelements, int Np, int dataKind, const double* coeffs, const double* x,
                                const double* y, const double* z, const double* wJ, double lambda, const int* mapB,
                                double* rhsL, double* ubL, void* stream) {
  const long NL = (long)Nelements * Np;
  if (NL <= 0) return LIBP_OK;
  if (dataKind == 1 && !coeffs) return arg_error("rhs_hex", "Fourier data needs coefficients");
  Data d{dataKind, coeffs};
  hipLaunchKernelGGL(rhs_kernel, dim3(stream_blocks(NL, 256)), dim3(256), 0, as_stream(stream), NL, d, x, y, z, wJ,
                     lambda, mapB, rhsL, ubL);
  return launch_status("rhs_hex");
}

extern "C" int libp_addbc_hex_f64(int NL, int dataKind, const double* coeffs, const double* x, const double* y,
                                  const double* z, const int* mapB, double* xL, void* stream) {
  if (NL <= 0) return LIBP_OK;
  Data d{dataKind, coeffs};
  hipLaunchKernelGGL(addbc_kernel, dim3(stream_blocks(NL, 256)), dim3(256), 0, as_stream(stream), (long)NL, d, x, y,
                     z, mapB, xL);
  return launch_status("addbc_hex");
}
