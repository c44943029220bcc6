// p-multigrid preconditioner (MultiGridPrecon, MGLevel, multigrid_t V-cycle)
// with the smoother and transfers in FP32 or FP64 and a Jacobi-PCG base level.
//
// Reference: solvers/elliptic/src/ellipticPreconMultiGrid.cpp:31-154,
// ellipticPreconMultiGridLevel.cpp:30-473, libs/parAlmond/parAlmondVcycle.cpp:34-60,
// and the multigridLevel ABC of include/parAlmond.hpp:70-91.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <random>
#include <type_traits>

#include "amg.hpp"
#include "elliptic.hpp"
#include "learned.hpp"

namespace libp {

/* Eigenvalues of a k x k upper Hessenberg matrix (column-major H[i+j*k]) by
 * complex single-shift QR with deflation; stands in for dgeev in
 * ellipticPreconMultiGridLevel.cpp:456-470 and parAlmondparCSR.cpp:444-446. */
std::vector<std::complex<double>> hessenbergEigenvalues(int n, const std::vector<double>& H) {
  typedef std::complex<double> cd;
  std::vector<cd> A(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) A[i * n + j] = H[i + j * n];
  double scale = 0;
  for (auto& a : A) scale = std::max(scale, std::abs(a));
  if (scale == 0) scale = 1;
  std::vector<cd> ev, gc(n), gs(n);
  int hi = n - 1, its = 0;
  while (hi >= 0) {
    if (hi == 0) {
      ev.push_back(A[0]);
      break;
    }
    int l = hi;
    for (; l > 0; --l) {
      double d = std::abs(A[(l - 1) * n + l - 1]) + std::abs(A[l * n + l]);
      if (d == 0) d = scale;
      if (std::abs(A[l * n + l - 1]) <= 1e-16 * d) {
        A[l * n + l - 1] = 0;
        break;
      }
    }
    if (l == hi) {
      ev.push_back(A[hi * n + hi]);
      --hi;
      its = 0;
      continue;
    }
    LIBP_ABORT("Hessenberg QR did not converge", ++its > 1000);
    const cd a = A[(hi - 1) * n + hi - 1], b = A[(hi - 1) * n + hi], c = A[hi * n + hi - 1], d = A[hi * n + hi];
    const cd m = 0.5 * (a + d), disc = std::sqrt(0.25 * (a - d) * (a - d) + b * c);
    cd mu = (std::abs(m + disc - d) < std::abs(m - disc - d)) ? m + disc : m - disc;
    if (its % 11 == 10) mu = d + std::abs(c);
    for (int i = l; i <= hi; ++i) A[i * n + i] -= mu;
    for (int i = l; i < hi; ++i) {
      const cd x = A[i * n + i], y = A[(i + 1) * n + i];
      const double r = std::sqrt(std::norm(x) + std::norm(y));
      gc[i] = (r == 0) ? cd(1) : x / r;
      gs[i] = (r == 0) ? cd(0) : y / r;
      for (int j = i; j <= hi; ++j) {
        const cd u = A[i * n + j], v = A[(i + 1) * n + j];
        A[i * n + j] = std::conj(gc[i]) * u + std::conj(gs[i]) * v;
        A[(i + 1) * n + j] = -gs[i] * u + gc[i] * v;
      }
    }
    for (int i = l; i < hi; ++i)
      for (int r = l; r <= std::min(i + 2, hi); ++r) {
        const cd u = A[r * n + i], v = A[r * n + i + 1];
        A[r * n + i] = u * gc[i] + v * gs[i];
        A[r * n + i + 1] = -u * std::conj(gs[i]) + v * std::conj(gc[i]);
      }
    for (int i = l; i <= hi; ++i) A[i * n + i] += mu;
  }
  return ev;
}

namespace {

template <typename T>
struct prec;
template <>
struct prec<double> {
  static int axpy(int N, double a, const double* x, double b, double* y, void* s) { return libp_axpy_f64(N, a, x, b, y, s); }
  static int amxpy(int N, double a, const double* A, const double* x, double b, double* y, void* s) {
    return libp_amxpy_f64(N, a, A, x, b, y, s);
  }
  static int coarsen(int Nf, int Nc, int E, const int* g2l, const double* P, const double* in, double* out, void* s) {
    return libp_mg_coarsen_hex_f64(Nf, Nc, E, nullptr, g2l, P, in, out, s);
  }
  static int prolong(int Nf, int Nc, int E, const int* g2l, const double* P, const double* in, double* out, void* s) {
    return libp_mg_prolong_hex_f64(Nf, Nc, E, nullptr, g2l, P, in, out, s);
  }
  static int chebStart(int N, const double* iD, const double* r, const double* Ax, double it, double* R, double* d,
                       double* X, int z, void* s) {
    return libp_cheb_start_f64(N, iD, r, Ax, it, R, d, X, z, s);
  }
  static int chebStep(int N, const double* iD, const double* Ad, double c1, double c2, double* R, double* d, double* X,
                      void* s) {
    return libp_cheb_step_f64(N, iD, Ad, c1, c2, R, d, X, s);
  }
  static void op(elliptic_t& e, deviceMemory<double>& x, deviceMemory<double>& y) { e.Operator(x, y); }
  static void res(elliptic_t& e, deviceMemory<double>& x, deviceMemory<double>& r, deviceMemory<double>& y) {
    e.OperatorResidual(x, r, y);
  }
  static int coarsenW(int Nf, int Nc, int E, const int* g2l, const double* P, const double* w, const double* in,
                      double* out, void* s) {
    return libp_mg_coarsen_weighted_hex_f64(Nf, Nc, E, nullptr, g2l, P, w, in, out, s);
  }
  static int gatherAdd(int N, const int* rs, const int* ci, const double* v, double* gv, void* s) {
    return libp_ogs_gather_add_f64(N, rs, ci, v, gv, s);
  }
  static const double* partialAx(elliptic_t& e, deviceMemory<double>& x) { return e.PartialAxLocal(x); }
  static int gatherChebStart(int N, const int* rs, const int* ci, const double* v, const double* iD, const double* r,
                             double it, double* R, double* d, double* X, void* s) {
    return libp_ogs_gather_cheb_start_f64(N, rs, ci, v, iD, r, it, R, d, X, s);
  }
  static int gatherChebStep(int N, const int* rs, const int* ci, const double* v, const double* iD, double c1,
                            double c2, double* R, double* d, double* X, void* s) {
    return libp_ogs_gather_cheb_step_f64(N, rs, ci, v, iD, c1, c2, R, d, X, s);
  }
  static deviceMemory<double> fromDouble(platform_t& p, deviceMemory<double>& v, dlong) { return v; }
};
template <>
struct prec<float> {
  static int axpy(int N, float a, const float* x, float b, float* y, void* s) { return libp_axpy_f32(N, a, x, b, y, s); }
  static int amxpy(int N, float a, const float* A, const float* x, float b, float* y, void* s) {
    return libp_amxpy_f32(N, a, A, x, b, y, s);
  }
  static int coarsen(int Nf, int Nc, int E, const int* g2l, const float* P, const float* in, float* out, void* s) {
    return libp_mg_coarsen_hex_f32(Nf, Nc, E, nullptr, g2l, P, in, out, s);
  }
  static int prolong(int Nf, int Nc, int E, const int* g2l, const float* P, const float* in, float* out, void* s) {
    return libp_mg_prolong_hex_f32(Nf, Nc, E, nullptr, g2l, P, in, out, s);
  }
  static int chebStart(int N, const float* iD, const float* r, const float* Ax, float it, float* R, float* d,
                       float* X, int z, void* s) {
    return libp_cheb_start_f32(N, iD, r, Ax, it, R, d, X, z, s);
  }
  static int chebStep(int N, const float* iD, const float* Ad, float c1, float c2, float* R, float* d, float* X,
                      void* s) {
    return libp_cheb_step_f32(N, iD, Ad, c1, c2, R, d, X, s);
  }
  static void op(elliptic_t& e, deviceMemory<float>& x, deviceMemory<float>& y) { e.OperatorF32(x, y); }
  static void res(elliptic_t& e, deviceMemory<float>& x, deviceMemory<float>& r, deviceMemory<float>& y) {
    e.OperatorResidualF32(x, r, y);
  }
  static int coarsenW(int Nf, int Nc, int E, const int* g2l, const float* P, const float* w, const float* in,
                      float* out, void* s) {
    return libp_mg_coarsen_weighted_hex_f32(Nf, Nc, E, nullptr, g2l, P, w, in, out, s);
  }
  static int gatherAdd(int N, const int* rs, const int* ci, const float* v, float* gv, void* s) {
    return libp_ogs_gather_add_f32(N, rs, ci, v, gv, s);
  }
  static const float* partialAx(elliptic_t& e, deviceMemory<float>& x) { return e.PartialAxLocalF32(x); }
  static int gatherChebStart(int N, const int* rs, const int* ci, const float* v, const float* iD, const float* r,
                             float it, float* R, float* d, float* X, void* s) {
    return libp_ogs_gather_cheb_start_f32(N, rs, ci, v, iD, r, it, R, d, X, s);
  }
  static int gatherChebStep(int N, const int* rs, const int* ci, const float* v, const float* iD, float c1, float c2,
                            float* R, float* d, float* X, void* s) {
    return libp_ogs_gather_cheb_step_f32(N, rs, ci, v, iD, c1, c2, R, d, X, s);
  }
  static deviceMemory<float> fromDouble(platform_t& p, deviceMemory<double>& v, dlong n) {
    deviceMemory<float> f = p.malloc<float>(n);
    LIBP_CALL(libp_convert_f64_to_f32(n, v.ptr(), f.ptr(), p.s()));
    return f;
  }
};

// Scratch shared by all levels (the reference keeps it static in MGLevel,
// ellipticPreconMultiGridLevel.cpp:208-221).
template <typename T>
struct mgScratch {
  deviceMemory<T> RES, Ad, d, transfer;
  void ensure(platform_t& p, dlong Ncols, dlong NL) {
    if ((dlong)RES.length() < Ncols) {
      RES = p.malloc<T>(Ncols);
      Ad = p.malloc<T>(Ncols);
      d = p.malloc<T>(Ncols);
    }
    if ((dlong)transfer.length() < NL) transfer = p.malloc<T>(NL);
  }
};

}  // namespace

namespace parAlmond {

/* include/parAlmond.hpp:70-91, templated on the preconditioner precision. */
template <typename T>
class multigridLevel {
 public:
  dlong Nrows = 0, Ncols = 0;
  virtual void Operator(deviceMemory<T>& o_X, deviceMemory<T>& o_Ax) = 0;
  virtual void smooth(deviceMemory<T>& o_RHS, deviceMemory<T>& o_X, bool x_is_zero) = 0;
  virtual void residual(deviceMemory<T>& o_RHS, deviceMemory<T>& o_X, deviceMemory<T>& o_RES) = 0;
  virtual void coarsen(deviceMemory<T>& o_X, deviceMemory<T>& o_Cx) = 0;
  virtual void prolongate(deviceMemory<T>& o_X, deviceMemory<T>& o_Px) = 0;
  virtual void Report() = 0;
  virtual ~multigridLevel() {}
};

}  // namespace parAlmond

template <typename T>
class MGLevel : public parAlmond::multigridLevel<T> {
 public:
  std::shared_ptr<elliptic_t> elliptic, ellipticC;
  platform_t platform;
  int Nf = 0, Nc = 0;
  enum { JACOBI = 0, CHEBYSHEV = 1, LEARNED = 2 } stype = CHEBYSHEV;
  int ChebyshevIterations = 2;
  double lambda0 = 0, lambda1 = 0;
  deviceMemory<double> o_invDiagA64;
  deviceMemory<T> o_invDiagA, o_P, o_weightG;
  std::shared_ptr<mgScratch<T>> scratch;
  std::shared_ptr<learnedSmoother_t> learned;  // fine-level learned smoother (FP32 only)
  // MULTIGRID LEARNED MODE: REPLACE (network pre- and post-smoothing), PRE (network
  // pre-smoothing, Chebyshev post) or AUGMENT (Chebyshev pre-smoothing followed by
  // one network step on its residual, Chebyshev post)
  enum { MODE_REPLACE = 0, MODE_PRE = 1, MODE_AUGMENT = 2 } learnedMode = MODE_REPLACE;
  std::vector<deviceMemory<T>>* record = nullptr;  // smoother input residuals are copied here when set
  size_t recordMax = 0;

  MGLevel(std::shared_ptr<elliptic_t> el, int _Nc, unsigned short* drandState, std::shared_ptr<mgScratch<T>> sc)
      : elliptic(el), platform(el->platform), Nf(el->mesh.N), Nc(_Nc), scratch(sc) {
    this->Nrows = el->Ndofs;
    this->Ncols = el->Ndofs + el->Nhalo;
    SetupSmoother(drandState);
    std::vector<double> P;
    degreeRaiseMatrix(Nc, Nf, P);
    deviceMemory<double> o_P64 = platform.malloc(P);
    o_P = prec<T>::fromDouble(platform, o_P64, (dlong)P.size());
    o_weightG = prec<T>::fromDouble(platform, el->o_weightG, el->Ndofs);
    scratch->ensure(platform, this->Ncols, el->mesh.Nelements * el->mesh.Np);
  }

  // MGLevel::SetupSmoother (ellipticPreconMultiGridLevel.cpp:345-384)
  void SetupSmoother(unsigned short* drandState) {
    const dlong N = elliptic->Ndofs;
    deviceMemory<double> o_diag = platform.malloc<double>(N);
    elliptic->BuildOperatorDiagonal(o_diag);
    std::vector<double> inv(N);
    o_diag.copyTo(inv.data(), N, 0, platform.stream);
    platform.finish();
    for (auto& v : inv) v = 1.0 / v;
    o_invDiagA64 = platform.malloc(inv);
    const settings_t& s = elliptic->settings;
    if (s.compareSetting("MULTIGRID SMOOTHER", "CHEBYSHEV") || s.compareSetting("MULTIGRID SMOOTHER", "LEARNED")) {
      stype = CHEBYSHEV;
      s.getSetting("MULTIGRID CHEBYSHEV DEGREE", ChebyshevIterations);
      const double rho = maxEigSmoothAx(drandState);
      lambda1 = rho;
      lambda0 = rho / 10.;
    } else {
      stype = JACOBI;
      const double rho = maxEigSmoothAx(drandState);
      lambda0 = (4. / 3.) / rho;
      for (auto& v : inv) v *= lambda0;
      o_invDiagA64.copyFrom(inv.data(), N, 0, platform.stream);
    }
    o_invDiagA = prec<T>::fromDouble(platform, o_invDiagA64, N);
    platform.finish();
  }

  // MGLevel::maxEigSmoothAx (ellipticPreconMultiGridLevel.cpp:393-473), in FP64.
  double maxEigSmoothAx(unsigned short* xs) {
    const dlong N = elliptic->Ndofs;
    int k = 10;
    if (k > N) k = N;
    std::vector<double> H(k * k, 0.0), Vx(N);
    std::vector<deviceMemory<double>> V(k + 1);
    for (int i = 0; i <= k; ++i) V[i] = platform.malloc<double>(N);
    deviceMemory<double> o_AVx = platform.malloc<double>(N);
    for (dlong i = 0; i < N; ++i) Vx[i] = erand48(xs);  // the drand48() stream seeded by srand48(0)
    deviceMemory<double> o_Vx = platform.malloc(Vx);
    void* st = platform.s();
    void* ws = platform.o_reduceWS.ptr();
    double* tmp = platform.o_scalars.ptr() + 12;
    LIBP_CALL(libp_norm2sq_f64(N, o_Vx.ptr(), tmp, ws, st));
    const double norm_vo = std::sqrt(platform.readScalars(tmp, 1)[0]);
    LIBP_CALL(libp_axpy_f64(N, 1. / norm_vo, o_Vx.ptr(), 0.0, V[0].ptr(), st));
    for (int j = 0; j < k; ++j) {
      elliptic->Operator(V[j], o_AVx);
      LIBP_CALL(libp_amxpy_f64(N, 1.0, o_invDiagA64.ptr(), o_AVx.ptr(), 0.0, V[j + 1].ptr(), st));
      for (int i = 0; i <= j; ++i) {
        LIBP_CALL(libp_inner_prod_f64(N, V[i].ptr(), V[j + 1].ptr(), tmp, ws, st));
        const double hij = platform.readScalars(tmp, 1)[0];
        LIBP_CALL(libp_axpy_f64(N, -hij, V[i].ptr(), 1.0, V[j + 1].ptr(), st));
        H[i + j * k] = hij;
      }
      if (j + 1 < k) {
        LIBP_CALL(libp_norm2sq_f64(N, V[j + 1].ptr(), tmp, ws, st));
        const double norm_vj = std::sqrt(platform.readScalars(tmp, 1)[0]);
        LIBP_CALL(libp_scale_f64(N, 1.0 / norm_vj, V[j + 1].ptr(), st));
        H[j + 1 + j * k] = norm_vj;
      }
    }
    double rho = 0;
    for (auto& l : hessenbergEigenvalues(k, H)) rho = std::max(rho, std::abs(l));
    return rho;
  }

  void Operator(deviceMemory<T>& o_X, deviceMemory<T>& o_Ax) override { prec<T>::op(*elliptic, o_X, o_Ax); }

  // RES = RHS - A X, the subtraction fused into the operator's gather
  void residual(deviceMemory<T>& o_RHS, deviceMemory<T>& o_X, deviceMemory<T>& o_RES) override {
    prec<T>::res(*elliptic, o_X, o_RHS, o_RES);
  }

  // MGLevel::coarsen (ellipticPreconMultiGridLevel.cpp:41-82); the inverse
  // multiplicity weight is applied as the kernel loads the fine residual
  void coarsen(deviceMemory<T>& o_X, deviceMemory<T>& o_Rx) override {
    deviceMemory<T>& o_RxL = scratch->transfer;
    LIBP_CALL(prec<T>::coarsenW(Nf, Nc, elliptic->mesh.Nelements, elliptic->o_GlobalToLocal.ptr(), o_P.ptr(),
                                o_weightG.ptr(), o_X.ptr(), o_RxL.ptr(), platform.s()));
    ellipticC->ogsMasked.Gather(o_Rx, o_RxL, 1, ogs::Add, ogs::Trans);
  }

  // MGLevel::prolongate (ellipticPreconMultiGridLevel.cpp:84-125); Px += the
  // gathered prolongation in the gather itself
  void prolongate(deviceMemory<T>& o_X, deviceMemory<T>& o_Px) override {
    deviceMemory<T>& o_PxL = scratch->transfer;
    LIBP_CALL(prec<T>::prolong(Nf, Nc, ellipticC->mesh.Nelements, ellipticC->o_GlobalToLocal.ptr(), o_P.ptr(),
                               o_X.ptr(), o_PxL.ptr(), platform.s()));
    const ogs::ogs_t& og = elliptic->ogsMasked;
    LIBP_CALL(prec<T>::gatherAdd(og.NlocalP, og.o_rowStartsN.ptr(), og.o_colIdsN.ptr(), o_PxL.ptr(), o_Px.ptr(),
                                 platform.s()));
  }

  bool learnedUsable() const { return learned && learned->active && !learned->fallback; }

  // copy of the residual the smoother is about to see (training data)
  void recordInput(deviceMemory<T>& o_RHS, deviceMemory<T>& o_X, bool x_is_zero) {
    if (record->size() >= recordMax) return;
    const dlong N = elliptic->Ndofs;
    deviceMemory<T> v = platform.malloc<T>(N);
    if (x_is_zero) {
      LIBP_HIP(hipMemcpyAsync(v.ptr(), o_RHS.ptr(), N * sizeof(T), hipMemcpyDeviceToDevice, platform.stream));
    } else {
      residual(o_RHS, o_X, v);
    }
    record->push_back(v);
  }

  // Pre-smoothing that also leaves the V-cycle residual in o_RES (the learned
  // step computes it on the way); false when the caller must do both itself.
  bool smoothPreResidual(deviceMemory<T>& o_RHS, deviceMemory<T>& o_X, deviceMemory<T>& o_RES) {
    if (learned && learnedMode == MODE_AUGMENT && (learnedUsable() || record)) {
      smoothChebyshev(o_RHS, o_X, true);
      residual(o_RHS, o_X, o_RES);
      if (record && record->size() < recordMax) {
        deviceMemory<T> v = platform.malloc<T>(elliptic->Ndofs);
        LIBP_HIP(hipMemcpyAsync(v.ptr(), o_RES.ptr(), elliptic->Ndofs * sizeof(T), hipMemcpyDeviceToDevice,
                                platform.stream));
        record->push_back(v);
      }
      if (learnedUsable()) augmentLearned(o_X, o_RES);
      return true;
    }
    if (!learnedUsable()) return false;
    if (record) recordInput(o_RHS, o_X, true);
    smoothLearned(o_RHS, o_X, true, &o_RES);
    return true;
  }

  // x += alpha c(res), res -= alpha A c (AUGMENT mode)
  void augmentLearned(deviceMemory<T>& o_X, deviceMemory<T>& o_RES);

  void smooth(deviceMemory<T>& o_RHS, deviceMemory<T>& o_X, bool x_is_zero) override {
    if (record && !(learnedMode == MODE_AUGMENT && learned)) recordInput(o_RHS, o_X, x_is_zero);
    if (learnedUsable() && (x_is_zero || learnedMode == MODE_REPLACE) && learnedMode != MODE_AUGMENT) {
      smoothLearned(o_RHS, o_X, x_is_zero, nullptr);
    } else if (stype == JACOBI) {
      smoothJacobi(o_RHS, o_X, x_is_zero);
    } else {
      smoothChebyshev(o_RHS, o_X, x_is_zero);
    }
  }

  void smoothJacobi(deviceMemory<T>& o_r, deviceMemory<T>& o_X, bool xIsZero) {
    const dlong N = elliptic->Ndofs;
    if (xIsZero) {
      LIBP_CALL(prec<T>::amxpy(N, T(1), o_invDiagA.ptr(), o_r.ptr(), T(0), o_X.ptr(), platform.s()));
      return;
    }
    deviceMemory<T>& o_RES = scratch->RES;
    Operator(o_X, o_RES);
    LIBP_CALL(prec<T>::axpy(N, T(1), o_r.ptr(), T(-1), o_RES.ptr(), platform.s()));
    LIBP_CALL(prec<T>::amxpy(N, T(1), o_invDiagA.ptr(), o_RES.ptr(), T(1), o_X.ptr(), platform.s()));
  }

  // MGLevel::smoothChebyshev (ellipticPreconMultiGridLevel.cpp:154-206) with
  // the vector updates fused: start = {RES, d, first x update}, each step =
  // {RES -= invD Ad, d = c1 RES + c2 d, x += d}.
  void smoothChebyshev(deviceMemory<T>& o_r, deviceMemory<T>& o_X, bool xIsZero) {
    const dlong N = elliptic->Ndofs;
    const double theta = 0.5 * (lambda1 + lambda0);
    const double delta = 0.5 * (lambda1 - lambda0);
    const double invTheta = 1.0 / theta;
    const double sigma = theta / delta;
    double rho_n = 1. / sigma, rho_np1;
    deviceMemory<T>& o_RES = scratch->RES;
    deviceMemory<T>& o_d = scratch->d;
    // each A d (and A x) is gathered and used at once: the Chebyshev update of a
    // row runs in the gather that forms its value (ogs_gather_cheb_kernel)
    const ogs::ogs_t& og = elliptic->ogsMasked;
    const int* rs = og.o_rowStartsT.ptr();
    const int* ci = og.o_colIdsT.ptr();
    if (xIsZero) {
      LIBP_CALL(prec<T>::chebStart(N, o_invDiagA.ptr(), o_r.ptr(), nullptr, (T)invTheta, o_RES.ptr(), o_d.ptr(),
                                   o_X.ptr(), 1, platform.s()));
    } else {
      const T* AxL = prec<T>::partialAx(*elliptic, o_X);
      LIBP_CALL(prec<T>::gatherChebStart(og.NlocalT, rs, ci, AxL, o_invDiagA.ptr(), o_r.ptr(), (T)invTheta,
                                         o_RES.ptr(), o_d.ptr(), o_X.ptr(), platform.s()));
    }
    for (int k = 0; k < ChebyshevIterations; k++) {
      const T* AdL = prec<T>::partialAx(*elliptic, o_d);
      rho_np1 = 1.0 / (2. * sigma - rho_n);
      const double rhoDivDelta = 2.0 * rho_np1 / delta;
      LIBP_CALL(prec<T>::gatherChebStep(og.NlocalT, rs, ci, AdL, o_invDiagA.ptr(), (T)rhoDivDelta,
                                        (T)(rho_np1 * rho_n), o_RES.ptr(), o_d.ptr(), o_X.ptr(), platform.s()));
      rho_n = rho_np1;
    }
  }

  void smoothLearned(deviceMemory<T>& o_r, deviceMemory<T>& o_X, bool xIsZero, deviceMemory<T>* o_RES);

  void Report() override {
    printf("|    pMG     |    %10lld  |   Matrix-free   |   %s|  Degree %2d\n", (long long)this->Nrows,
           stype == CHEBYSHEV ? "Chebyshev       " : "Damped Jacobi   ", Nf);
  }
};

template <>
void MGLevel<float>::smoothLearned(deviceMemory<float>& o_r, deviceMemory<float>& o_X, bool xIsZero,
                                   deviceMemory<float>* o_RES) {
  auto A = [this](deviceMemory<float>& x, deviceMemory<float>& y) { Operator(x, y); };
  auto R = [this](deviceMemory<float>& x, deviceMemory<float>& r, deviceMemory<float>& res) { residual(r, x, res); };
  if (xIsZero) {
    learned->smoothPre(o_r, o_X, o_RES, o_invDiagA.ptr(), A, R);
  } else {
    learned->smoothPost(o_r, o_X, o_invDiagA.ptr(), A, R);
  }
}
template <>
void MGLevel<double>::smoothLearned(deviceMemory<double>&, deviceMemory<double>&, bool, deviceMemory<double>*) {
  LIBP_FORCE_ABORT("the learned smoother runs in FP32 only");
}
template <>
void MGLevel<float>::augmentLearned(deviceMemory<float>& o_X, deviceMemory<float>& o_RES) {
  auto A = [this](deviceMemory<float>& x, deviceMemory<float>& y) { Operator(x, y); };
  learned->augment(o_X, o_RES, o_invDiagA.ptr(), A);
}
template <>
void MGLevel<double>::augmentLearned(deviceMemory<double>&, deviceMemory<double>&) {
  LIBP_FORCE_ABORT("the learned smoother runs in FP32 only");
}

/* Base level: FP64 Jacobi-PCG on the N=1 operator (the composition the p-MG
 * configuration asks for; the reference ends in parAlmond AMG instead,
 * ellipticPreconMultiGrid.cpp:133-153). Tolerance is relative to the
 * incoming residual; a zero tolerance runs MAXIT iterations. */
class coarseJacobiPCG_t {
 public:
  std::shared_ptr<elliptic_t> elliptic;
  std::shared_ptr<amgSolver_t> amg;  // [MULTIGRID COARSE SOLVER] AMG: one V-cycle replaces the PCG
  std::shared_ptr<LinearSolver::pcg> solver;
  std::shared_ptr<JacobiPrecon> jacobi;
  deviceMemory<double> o_x, o_r;
  double tol = 1e-2;
  int maxit = 200;
  int lastIterations = 0;
  coarseJacobiPCG_t(std::shared_ptr<elliptic_t> el, unsigned short xs[3]) : elliptic(el) {
    elliptic->settings.getSetting("MULTIGRID COARSE TOLERANCE", tol);
    elliptic->settings.getSetting("MULTIGRID COARSE MAXIT", maxit);
    jacobi = std::make_shared<JacobiPrecon>(*elliptic);
    settings_t s = elliptic->settings;
    s.changeSetting("LINEAR SOLVER", "PCG");
    s.changeSetting("LINEAR SOLVER STOPPING CRITERION", "RELATIVE");
    solver = std::make_shared<LinearSolver::pcg>(elliptic->Ndofs, 0, elliptic->platform, s, elliptic->comm);
    solver->pureRelative = true;
    elliptic->settings.getSetting("MULTIGRID COARSE CHECK EVERY", solver->checkEvery);
    o_x = elliptic->platform.malloc<double>(elliptic->Ndofs);
    o_r = elliptic->platform.malloc<double>(elliptic->Ndofs);
    // fused three-kernel iterations on an N=1 level (not with the all-Neumann mean removal)
    std::string fused = "TRUE";
    elliptic->settings.getSetting("MULTIGRID COARSE FUSED", fused);
    fusedHex1 = elliptic->mesh.N == 1 && !elliptic->allNeumann && fused == "TRUE";
    std::string graph = "TRUE";
    elliptic->settings.getSetting("MULTIGRID COARSE GRAPH", graph);
    useGraph = graph == "TRUE";
    // one fused reduction per iteration (Chronopoulos-Gear), two launches
    std::string cg1s = "TRUE";
    elliptic->settings.getSetting("MULTIGRID COARSE SINGLE REDUCTION", cg1s);
    cg1 = fusedHex1 && cg1s == "TRUE";
    // single-reduction iterations on the assembled N=1 operator: one launch each
    if (cg1 && elliptic->settings.compareSetting("MULTIGRID COARSE ASSEMBLED", "TRUE") &&
        !elliptic->settings.compareSetting("MULTIGRID COARSE SOLVER", "AMG")) {
      const csrHost_t A = ellipticAssembleCSR(*elliptic);
      LIBP_ABORT("assembled coarse operator: size mismatch", A.n != elliptic->Ndofs || A.m != elliptic->Ndofs);
      platform_t& p = elliptic->platform;
      o_Crs = p.malloc(A.rs);
      o_Cci = p.malloc(A.ci);
      o_Cv = p.malloc(A.v);
      cg1csr = true;
    }
    // the reference's base level: parAlmond AMG on the assembled operator
    // (ellipticPreconMultiGrid.cpp:119-150), one V-cycle per coarse solve; on the
    // all-Neumann operator the coarsest matrix carries the null-space penalty
    if (elliptic->settings.compareSetting("MULTIGRID COARSE SOLVER", "AMG"))
      amg = std::make_shared<amgSolver_t>(*elliptic, xs);
  }
  ~coarseJacobiPCG_t() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    for (int k = 0; k < 2; ++k) {
      if (h_pipe[k]) (void)hipHostFree(h_pipe[k]);
      if (ev_pipe[k]) (void)hipEventDestroy(ev_pipe[k]);
    }
  }
  bool fusedHex1 = false;
  bool useGraph = true;
  bool cg1 = false;
  bool cg1csr = false;  // cg1 on the assembled operator (o_Crs, o_Cci, o_Cv; w in o_w, o_w1)
  deviceMemory<dlong> o_Crs, o_Cci;
  deviceMemory<double> o_Cv, o_w1;
  deviceMemory<char> o_ws2;  // two reduction workspaces (the assembled form alternates them)
  deviceMemory<double> o_s, o_z, o_p, o_Ap, o_AqL;
  // single-reduction form: w, the parity pairs of p, s, r (r[0] is o_r), u, owner flags
  deviceMemory<double> o_w, o_p2[2], o_s2[2], o_r1, o_u2[2];
  deviceMemory<unsigned char> o_owner;
  // captured chunks of fused iterations, per solution vector and first-iteration flag
  struct chunkGraph {
    double* x;
    int first, n;
    hipGraphExec_t exec;
    int kind = 0, parity = 0;   // kind 1: single-reduction chunk starting at this parity
    double* mirror = nullptr;  // host mirror its last check writes (assembled form)
  };
  std::vector<chunkGraph> graphs;
  // pinned, device-mapped copies of the scalars after each of two chunks in
  // flight (host and device pointers), and their events
  double* h_pipe[2] = {nullptr, nullptr};
  double* d_pipe[2] = {nullptr, nullptr};
  hipEvent_t ev_pipe[2] = {nullptr, nullptr};
  template <class Enqueue>
  int runChunks(int n0, Enqueue enqueue);
  void enqueueChunk(deviceMemory<double>& o_X, int firstIt, int n);
  int solveFusedHex1(deviceMemory<double>& o_X);
  // returns true when the chunk's last check wrote the scalars to mirror itself
  bool enqueueCG1Chunk(deviceMemory<double>& o_X, int itStart, int n, double* mirror);
  int solveCG1Hex1(deviceMemory<double>& o_X);
  void allocFused();
  // iterations of the last coarse solve
  int iterations() const { return lastIterations; }
  template <typename T>
  void solve(deviceMemory<T>& o_rhs, deviceMemory<T>& o_X);
};

/* Jacobi-PCG from x = 0 with the N=1 operator fused into three kernels per
 * iteration (libp_pcg_jacobi_hex1_iteration_f64): the same iterates and
 * stopping rule as LinearSolver::pcg with JacobiPrecon on this level. */
void coarseJacobiPCG_t::allocFused() {
  if (o_s.ptr()) return;
  platform_t& p = elliptic->platform;
  const dlong N = elliptic->Ndofs;
  const mesh_t& m = elliptic->mesh;
  o_s = p.malloc<double>(LIBP_PCG_NSCALARS);
  o_z = p.malloc<double>(N);
  o_p = p.malloc<double>(N);
  o_Ap = p.malloc<double>(N);
  o_AqL = p.malloc<double>((size_t)m.Nelements * m.Np);
}

/* Queues chunks of iterations until the device STOP flag or maxit. The first
 * chunk has n0 iterations, the rest min(checkEvery, 4). After each chunk a
 * kernel writes the scalars into pinned host memory, behind an event, and the
 * next chunk is queued before the host waits for the previous event, so the
 * GPU does not sit idle while the host reads the flag.
 * Iterations queued past convergence return at once. Returns the iteration
 * count of the device. */
template <class Enqueue>
int coarseJacobiPCG_t::runChunks(int n0, Enqueue enqueue) {
  platform_t& p = elliptic->platform;
  if (maxit <= 0) return 0;
  if (!h_pipe[0])
    for (int k = 0; k < 2; ++k) {
      LIBP_HIP(hipHostMalloc((void**)&h_pipe[k], LIBP_PCG_NSCALARS * sizeof(double),
                             hipHostMallocMapped | hipHostMallocCoherent));
      LIBP_HIP(hipHostGetDevicePointer((void**)&d_pipe[k], h_pipe[k], 0));
      LIBP_HIP(hipEventCreateWithFlags(&ev_pipe[k], hipEventDisableTiming));
    }
  const int step = std::min(std::max(1, solver->checkEvery), 4);
  int queued = 0, slot = 0, it = 0;
  auto launch = [&](int n, int sl) {
    const bool mirrored = enqueue(queued, n, d_pipe[sl]);
    queued += n;
    if (!mirrored) LIBP_CALL(libp_scalars_to_host_f64(LIBP_PCG_NSCALARS, o_s.ptr(), d_pipe[sl], p.s()));
    LIBP_HIP(hipEventRecord(ev_pipe[sl], p.stream));
  };
  launch(std::max(1, std::min(n0, maxit)), 0);
  for (;;) {
    const int n = std::min(step, maxit - queued);
    if (n > 0) launch(n, slot ^ 1);
    LIBP_HIP(hipEventSynchronize(ev_pipe[slot]));
    it = (int)h_pipe[slot][LIBP_PCG_ITERS];
    if (h_pipe[slot][LIBP_PCG_STOP] != 0.0 || n <= 0) break;
    slot ^= 1;
  }
  return it;
}

int coarseJacobiPCG_t::solveFusedHex1(deviceMemory<double>& o_X) {
  platform_t& p = elliptic->platform;
  void* st = p.s();
  void* ws = p.o_reduceWS.ptr();
  const dlong N = elliptic->Ndofs;
  allocFused();
  double* s = o_s.ptr();
  LIBP_CALL(libp_set_f64(N, 0.0, o_X.ptr(), st));
  LIBP_CALL(libp_set_f64(LIBP_PCG_NSCALARS, 0.0, s, st));
  LIBP_CALL(libp_norm2sq_f64(N, o_r.ptr(), s + LIBP_PCG_RDOTR, ws, st));
  // TOL = tol^2 r0.r0 on the device; a zero residual raises STOP at once
  LIBP_CALL(libp_pcg_relative_tol_f64(tol, s, st));
  LIBP_CALL(libp_pcg_jacobi_rdotz_f64(N, jacobi->o_invDiagA.ptr(), o_r.ptr(), o_z.ptr(), o_Ap.ptr(), 0, s, ws, st));
  // The first chunk is sized from the previous solve (the coarse problem is
  // solved once per V-cycle, with similar counts), rounded down to a multiple
  // of 4 so that few distinct chunks are captured; the rest go in chunks of
  // 4. Iterations past convergence return at once (device STOP flag).
  const int chunk = std::max(1, solver->checkEvery);
  const int n0 = (lastIterations >= 8) ? (lastIterations / 4) * 4 : chunk;
  return runChunks(n0, [&](int itStart, int n, double*) {
    enqueueChunk(o_X, itStart == 0, n);
    return false;
  });
}

/* n fused iterations (the first with first = 1 when firstIt). With
 * [MULTIGRID COARSE GRAPH] TRUE the chunk is captured once as a HIP graph and
 * replayed: the kernels take only fixed device pointers, read their scalars
 * from device memory and stop themselves, so a replay is the same work as the
 * launches it records (no memset nodes: the reduction counters reset
 * themselves). */
void coarseJacobiPCG_t::enqueueChunk(deviceMemory<double>& o_X, int firstIt, int n) {
  platform_t& p = elliptic->platform;
  void* st = p.s();
  void* ws = p.o_reduceWS.ptr();
  const mesh_t& m = elliptic->mesh;
  const dlong N = elliptic->Ndofs;
  auto launch = [&]() {
    for (int c = 0; c < n; ++c)
      LIBP_CALL(libp_pcg_jacobi_hex1_iteration_f64(
          m.Nelements, elliptic->o_GlobalToLocal.ptr(), m.o_wJ.ptr(), m.o_ggeo.ptr(), m.o_D.ptr(), elliptic->lambda,
          N, elliptic->ogsMasked.o_rowStartsT.ptr(), elliptic->ogsMasked.o_colIdsT.ptr(), jacobi->o_invDiagA.ptr(),
          firstIt && c == 0, o_X.ptr(), o_r.ptr(), o_z.ptr(), o_p.ptr(), o_Ap.ptr(), o_AqL.ptr(), o_s.ptr(), ws, st));
    // the stopping test of the last iteration, before the host reads the flag
    LIBP_CALL(libp_pcg_jacobi_hex1_check_f64(N, o_s.ptr(), ws, st));
  };
  if (!useGraph) {
    launch();
    return;
  }
  for (auto& g : graphs)
    if (g.kind == 0 && g.x == o_X.ptr() && g.first == firstIt && g.n == n) {
      LIBP_HIP(hipGraphLaunch(g.exec, p.stream));
      return;
    }
  hipGraph_t graph;
  LIBP_HIP(hipStreamBeginCapture(p.stream, hipStreamCaptureModeThreadLocal));
  launch();
  LIBP_HIP(hipStreamEndCapture(p.stream, &graph));
  chunkGraph g{o_X.ptr(), firstIt, n, nullptr, 0, 0};
  LIBP_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
  LIBP_HIP(hipGraphDestroy(graph));
  graphs.push_back(g);
  LIBP_HIP(hipGraphLaunch(g.exec, p.stream));
}

/* The fused Jacobi-PCG in its single-reduction form
 * (libp_pcg_jacobi_hex1_cg1_f64): two launches per iteration, the PCG iterates
 * in exact arithmetic, the same stopping rule. On entry o_r holds the
 * right-hand side; it is the parity-0 residual buffer. */
int coarseJacobiPCG_t::solveCG1Hex1(deviceMemory<double>& o_X) {
  platform_t& p = elliptic->platform;
  const dlong N = elliptic->Ndofs;
  const mesh_t& m = elliptic->mesh;
  allocFused();
  if (!o_w.ptr()) {
    o_w = p.malloc<double>(N);
    for (int k = 0; k < 2; ++k) {
      o_p2[k] = p.malloc<double>(N);
      o_s2[k] = p.malloc<double>(N);
      o_u2[k] = p.malloc<double>(N);
    }
    o_r1 = p.malloc<double>(N);
    if (cg1csr) {
      o_w1 = p.malloc<double>(N);
      o_ws2 = deviceMemory<char>::alloc(2 * (size_t)libp_reduce_ws_bytes());
    }
    // the first local copy of each gathered node writes its updates
    const ogs::ogs_t& og = elliptic->ogsMasked;
    std::vector<unsigned char> own((size_t)m.Nelements * m.Np, 0);
    for (dlong g = 0; g < og.NlocalT; ++g)
      if (og.rowStartsT[g + 1] > og.rowStartsT[g]) own[og.colIdsT[og.rowStartsT[g]]] = 1;
    o_owner = p.malloc(own);
  }
  LIBP_CALL(libp_set_f64(LIBP_PCG_NSCALARS, 0.0, o_s.ptr(), p.s()));
  // first chunk sized from the previous solve, as in solveFusedHex1
  const int chunk = std::max(1, solver->checkEvery);
  const int n0 = (lastIterations >= 8) ? (lastIterations / 4) * 4 : chunk;
  return runChunks(n0, [&](int itStart, int n, double* mirror) { return enqueueCG1Chunk(o_X, itStart, n, mirror); });
}

bool coarseJacobiPCG_t::enqueueCG1Chunk(deviceMemory<double>& o_X, int itStart, int n, double* mirror) {
  platform_t& p = elliptic->platform;
  void* st = p.s();
  void* ws = p.o_reduceWS.ptr();
  const mesh_t& m = elliptic->mesh;
  const dlong N = elliptic->Ndofs;
  double* P2[2] = {o_p2[0].ptr(), o_p2[1].ptr()};
  double* S2[2] = {o_s2[0].ptr(), o_s2[1].ptr()};
  double* R2[2] = {o_r.ptr(), o_r1.ptr()};
  double* U2[2] = {o_u2[0].ptr(), o_u2[1].ptr()};
  const int first = itStart == 0;
  double* W2[2] = {o_w.ptr(), o_w1.ptr()};
  auto call = [&](int mode, int parity) {
    if (cg1csr) {
      LIBP_CALL(libp_pcg_jacobi_csr_cg1_f64(N, o_Crs.ptr(), o_Cci.ptr(), o_Cv.ptr(), jacobi->o_invDiagA.ptr(), tol,
                                            mode, parity, o_X.ptr(), W2, P2, S2, R2, U2, o_s.ptr(), o_ws2.ptr(), st));
      return;
    }
    LIBP_CALL(libp_pcg_jacobi_hex1_cg1_f64(
        m.Nelements, elliptic->o_GlobalToLocal.ptr(), o_owner.ptr(), m.o_wJ.ptr(), m.o_ggeo.ptr(), m.o_D.ptr(),
        elliptic->lambda, N, elliptic->ogsMasked.o_rowStartsT.ptr(), elliptic->ogsMasked.o_colIdsT.ptr(),
        jacobi->o_invDiagA.ptr(), tol, mode, parity, o_X.ptr(), o_w.ptr(), P2, S2, R2, U2, o_AqL.ptr(), o_s.ptr(), ws,
        st));
  };
  // the assembled form's last check also writes the scalars to the host mirror
  double* const hostMirror = cg1csr ? mirror : nullptr;
  auto launch = [&]() {
    if (first) call(0, 0);
    for (int c = 0; c < n; ++c) {
      const int itc = itStart + c;
      call(itc == 0 ? 1 : 2, itc & 1);
    }
    // the stopping test of the last iteration (of this parity)
    if (hostMirror)
      LIBP_CALL(libp_pcg_jacobi_csr_cg1_check_f64(N, (itStart + n - 1) & 1, o_s.ptr(), o_ws2.ptr(), hostMirror, st));
    else
      call(3, (itStart + n - 1) & 1);
  };
  if (!useGraph) {
    launch();
    return hostMirror != nullptr;
  }
  const int parity = itStart & 1;
  for (auto& g : graphs)
    if (g.kind == 1 && g.x == o_X.ptr() && g.first == first && g.n == n && g.parity == parity &&
        g.mirror == hostMirror) {
      LIBP_HIP(hipGraphLaunch(g.exec, p.stream));
      return hostMirror != nullptr;
    }
  hipGraph_t graph;
  LIBP_HIP(hipStreamBeginCapture(p.stream, hipStreamCaptureModeThreadLocal));
  launch();
  LIBP_HIP(hipStreamEndCapture(p.stream, &graph));
  chunkGraph g{o_X.ptr(), first, n, nullptr, 1, parity, hostMirror};
  LIBP_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
  LIBP_HIP(hipGraphDestroy(graph));
  graphs.push_back(g);
  LIBP_HIP(hipGraphLaunch(g.exec, p.stream));
  return hostMirror != nullptr;
}

template <>
void coarseJacobiPCG_t::solve<double>(deviceMemory<double>& o_rhs, deviceMemory<double>& o_X) {
  platform_t& p = elliptic->platform;
  const dlong N = elliptic->Ndofs;
  if (amg) {
    amg->vcycle(o_rhs.ptr(), o_X.ptr());
    lastIterations = 1;
    return;
  }
  LIBP_HIP(hipMemcpyAsync(o_r.ptr(), o_rhs.ptr(), N * sizeof(double), hipMemcpyDeviceToDevice, p.stream));
  if (cg1) {
    lastIterations = solveCG1Hex1(o_X);
    return;
  }
  if (fusedHex1) {
    lastIterations = solveFusedHex1(o_X);
    return;
  }
  LIBP_CALL(libp_set_f64(N, 0.0, o_X.ptr(), p.s()));
  lastIterations = solver->Solve(*elliptic, *jacobi, o_X, o_r, tol, maxit, 0);
}
template <>
void coarseJacobiPCG_t::solve<float>(deviceMemory<float>& o_rhs, deviceMemory<float>& o_X) {
  platform_t& p = elliptic->platform;
  const dlong N = elliptic->Ndofs;
  LIBP_CALL(libp_convert_f32_to_f64(N, o_rhs.ptr(), o_r.ptr(), p.s()));
  if (amg) {
    amg->vcycle(o_r.ptr(), o_x.ptr());
    lastIterations = 1;
  } else if (cg1) {
    lastIterations = solveCG1Hex1(o_x);
  } else if (fusedHex1) {
    lastIterations = solveFusedHex1(o_x);
  } else {
    LIBP_CALL(libp_set_f64(N, 0.0, o_x.ptr(), p.s()));
    lastIterations = solver->Solve(*elliptic, *jacobi, o_x, o_r, tol, maxit, 0);
  }
  LIBP_CALL(libp_convert_f64_to_f32(N, o_x.ptr(), o_X.ptr(), p.s()));
}

class MultiGridPreconBase {
 public:
  virtual void Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) = 0;
  virtual bool OperatorRdotz(deviceMemory<dfloat>&, deviceMemory<dfloat>&, deviceMemory<dfloat>&, int, double*,
                             void*) {
    return false;
  }
  virtual ~MultiGridPreconBase() {}
};

template <typename T>
class multigrid_t : public MultiGridPreconBase {
 public:
  elliptic_t& fine;
  std::vector<std::shared_ptr<elliptic_t>> ellipticLevels;
  std::vector<std::shared_ptr<MGLevel<T>>> levels;
  std::shared_ptr<coarseJacobiPCG_t> coarse;
  std::vector<deviceMemory<T>> o_rhs, o_x;
  deviceMemory<T> o_res, o_r0, o_z0;
  std::shared_ptr<mgScratch<T>> scratch = std::make_shared<mgScratch<T>>();

  explicit multigrid_t(elliptic_t& _fine) : fine(_fine) {
    std::vector<int> deg = multigridDegrees(fine.settings, fine.mesh.N);
    LIBP_ABORT("p-multigrid needs at least two degrees", deg.size() < 2);
    unsigned short xs[3] = {0x330E, 0, 0};  // srand48(0): parAlmondKernels.cpp:55-58 on rank 0
    auto el0 = std::make_shared<elliptic_t>(fine);  // SetupNewDegree(N) returns the fine operator
    ellipticLevels.push_back(el0);
    for (size_t l = 0; l + 1 < deg.size(); ++l) {
      auto lev = std::make_shared<MGLevel<T>>(ellipticLevels[l], deg[l + 1], xs, scratch);
      levels.push_back(lev);
      mesh_t meshC = fine.mesh.SetupNewDegree(deg[l + 1]);
      auto elC = std::make_shared<elliptic_t>(fine.SetupNewDegree(meshC));
      ellipticLevels.push_back(elC);
      lev->ellipticC = elC;
    }
    coarse = std::make_shared<coarseJacobiPCG_t>(ellipticLevels.back(), xs);
    o_rhs.resize(ellipticLevels.size());
    o_x.resize(ellipticLevels.size());
    dlong maxN = 0;
    for (size_t l = 0; l < ellipticLevels.size(); ++l) {
      const dlong N = ellipticLevels[l]->Ndofs;
      maxN = std::max(maxN, N);
      if (l > 0) {
        o_rhs[l] = fine.platform.malloc<T>(N);
        o_x[l] = fine.platform.malloc<T>(N);
      }
    }
    o_res = fine.platform.malloc<T>(maxN);
    if (std::is_same<T, float>::value) {
      o_r0 = fine.platform.malloc<T>(fine.Ndofs);
      o_z0 = fine.platform.malloc<T>(fine.Ndofs);
    }
    if (fine.settings.compareSetting("MULTIGRID SMOOTHER", "LEARNED")) attachLearned();
    if (fine.settings.compareSetting("VERBOSE", "TRUE")) {
      printf("p-multigrid levels (%s):\n", std::is_same<T, float>::value ? "FP32" : "FP64");
      for (auto& l : levels) l->Report();
      printf("|   coarse   |    %10lld  |  Jacobi-PCG FP64 |  Degree %2d\n", (long long)ellipticLevels.back()->Ndofs,
             ellipticLevels.back()->mesh.N);
    }
  }

  void attachLearned();
  std::vector<double> trainLearned(const learnedTrainOptions& opt, int verbose);
  std::vector<double> trainLinear(const learnedTrainOptions& opt, int verbose);

  // parAlmondVcycle.cpp:34-60
  void vcycle(size_t k, deviceMemory<T>& o_RHS, deviceMemory<T>& o_X) {
    if (k + 1 == ellipticLevels.size()) {
      coarse->solve<T>(o_RHS, o_X);
      return;
    }
    MGLevel<T>& level = *levels[k];
    if (!level.smoothPreResidual(o_RHS, o_X, o_res)) {
      level.smooth(o_RHS, o_X, true);
      level.residual(o_RHS, o_X, o_res);
    }
    level.coarsen(o_res, o_rhs[k + 1]);
    vcycle(k + 1, o_rhs[k + 1], o_x[k + 1]);
    level.prolongate(o_x[k + 1], o_X);
    level.smooth(o_RHS, o_X, false);
  }

  void Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) override;
  bool OperatorRdotz(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_z, deviceMemory<dfloat>& o_Ap, int flexible,
                     double* scalars, void* ws) override;
};

template <>
bool multigrid_t<double>::OperatorRdotz(deviceMemory<dfloat>&, deviceMemory<dfloat>&, deviceMemory<dfloat>&, int,
                                        double*, void*) {
  return false;
}
// the V-cycle's FP32 output widened to z in the pass that forms r.z and z.Ap
template <>
bool multigrid_t<float>::OperatorRdotz(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_z,
                                       deviceMemory<dfloat>& o_Ap, int flexible, double* scalars, void* ws) {
  const dlong N = fine.Ndofs;
  LIBP_CALL(libp_convert_f64_to_f32(N, o_r.ptr(), o_r0.ptr(), fine.platform.s()));
  vcycle(0, o_r0, o_z0);
  LIBP_CALL(libp_pcg_rdotz_from_f32_f64(N, o_r.ptr(), o_z0.ptr(), o_z.ptr(), o_Ap.ptr(), flexible, scalars, ws,
                                        fine.platform.s()));
  return true;
}

template <>
void multigrid_t<double>::Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) {
  vcycle(0, o_r, o_Mr);
}
template <>
void multigrid_t<float>::Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) {
  const dlong N = fine.Ndofs;
  LIBP_CALL(libp_convert_f64_to_f32(N, o_r.ptr(), o_r0.ptr(), fine.platform.s()));
  vcycle(0, o_r0, o_z0);
  LIBP_CALL(libp_convert_f32_to_f64(N, o_z0.ptr(), o_Mr.ptr(), fine.platform.s()));
}

template <>
void multigrid_t<float>::attachLearned() {
  std::string path = fine.settings.getSetting("MULTIGRID LEARNED WEIGHTS");
  const int arch = fine.settings.compareSetting("MULTIGRID LEARNED ARCH", "LINEAR") ? learnedSmoother_t::ARCH_LINEAR
                                                                                     : learnedSmoother_t::ARCH_MLP;
  auto ls = std::make_shared<learnedSmoother_t>(*ellipticLevels[0], arch);
  fine.settings.getSetting("MULTIGRID LEARNED SWEEPS", ls->sweeps);
  LIBP_ABORT("MULTIGRID LEARNED SWEEPS must be at least 1", ls->sweeps < 1);
  if (!path.empty() && !ls->load(path) && fine.settings.compareSetting("VERBOSE", "TRUE"))
    printf("learned smoother %s; smoothing with Chebyshev\n", ls->status.c_str());
  levels[0]->learned = ls;
  if (fine.settings.compareSetting("MULTIGRID LEARNED MODE", "PRE")) levels[0]->learnedMode = MGLevel<float>::MODE_PRE;
  if (fine.settings.compareSetting("MULTIGRID LEARNED MODE", "AUGMENT"))
    levels[0]->learnedMode = MGLevel<float>::MODE_AUGMENT;
}
template <>
void multigrid_t<double>::attachLearned() {
  LIBP_FORCE_ABORT("MULTIGRID SMOOTHER LEARNED needs MULTIGRID PRECISION FP32");
}

MultiGridPrecon::MultiGridPrecon(elliptic_t& _elliptic) : elliptic(_elliptic) {
  if (elliptic.settings.compareSetting("MULTIGRID PRECISION", "FP64"))
    impl = std::make_shared<multigrid_t<double>>(elliptic);
  else
    impl = std::make_shared<multigrid_t<float>>(elliptic);
}

// MultiGridPrecon::Operator (ellipticPreconMultiGrid.cpp:31-38)
void MultiGridPrecon::Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) {
  impl->Operator(o_r, o_Mr);
  if (elliptic.allNeumann) elliptic.ZeroMean(o_Mr);
}

bool MultiGridPrecon::OperatorRdotz(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_z, deviceMemory<dfloat>& o_Ap,
                                    int flexible, double* scalars, void* ws) {
  if (elliptic.allNeumann) return false;  // the mean is removed after the V-cycle
  return impl->OperatorRdotz(o_r, o_z, o_Ap, flexible, scalars, ws);
}

void MultiGridPrecon::solveStarting() {
  if (learnedSmoother_t* L = multigridLearned(*this)) L->monitorStart();
}

void MultiGridPrecon::iterationDone(int it, double rdotr) {
  if (learnedSmoother_t* L = multigridLearned(*this)) L->monitorIteration(it, rdotr);
}

namespace {

// FPCG preconditioner: the FP32 V-cycle started at level k of a multigrid.
class levelVcycle_t : public operator_t {
 public:
  multigrid_t<float>& mg;
  size_t k;
  levelVcycle_t(multigrid_t<float>& m, size_t kk) : mg(m), k(kk) {}
  void Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_z) override {
    if (k == 0) {
      mg.Operator(o_r, o_z);
      return;
    }
    void* s = mg.fine.platform.s();
    const dlong N = mg.ellipticLevels[k]->Ndofs;
    LIBP_CALL(libp_convert_f64_to_f32(N, o_r.ptr(), mg.o_rhs[k].ptr(), s));
    mg.vcycle(k, mg.o_rhs[k], mg.o_x[k]);
    LIBP_CALL(libp_convert_f32_to_f64(N, mg.o_x[k].ptr(), o_z.ptr(), s));
  }
};

// x = A^{-1} b by FPCG from a zero guess to a relative tolerance; b is kept
int fpcgSolve(elliptic_t& el, operator_t& M, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_b,
              deviceMemory<dfloat>& o_work, double tol, int maxit) {
  settings_t s = el.settings;
  s.changeSetting("LINEAR SOLVER", "FPCG");
  LinearSolver::pcg solver(el.Ndofs, el.Nhalo, el.platform, s, el.comm);
  solver.pureRelative = true;
  LIBP_CALL(libp_set_f64(el.Ndofs, 0.0, o_x.ptr(), el.platform.s()));
  LIBP_HIP(hipMemcpyAsync(o_work.ptr(), o_b.ptr(), el.Ndofs * sizeof(double), hipMemcpyDeviceToDevice,
                          el.platform.stream));
  return solver.Solve(el, M, o_x, o_work, tol, maxit, 0);
}

}  // namespace

template <>
std::vector<double> multigrid_t<float>::trainLinear(const learnedTrainOptions& opt, int verbose);

template <>
std::vector<double> multigrid_t<float>::trainLearned(const learnedTrainOptions& opt, int verbose) {
  LIBP_ABORT("training needs MULTIGRID SMOOTHER LEARNED", levels.empty() || !levels[0]->learned);
  LIBP_ABORT("training needs at least two p-MG levels above the coarse solve", ellipticLevels.size() < 3);
  learnedSmoother_t& L = *levels[0]->learned;
  MGLevel<float>& lev0 = *levels[0];
  elliptic_t& el0 = *ellipticLevels[0];
  elliptic_t& el1 = *ellipticLevels[1];
  platform_t& p = fine.platform;
  void* st = p.s();
  const dlong N0 = el0.Ndofs, N1 = el1.Ndofs;
  levelVcycle_t M0(*this, 0), M1(*this, 1);
  if (L.arch == learnedSmoother_t::ARCH_LINEAR) return trainLinear(opt, verbose);
  deviceMemory<double> b = p.malloc<double>(N0), x = p.malloc<double>(N0), w0 = p.malloc<double>(N0);
  deviceMemory<double> r64 = p.malloc<double>(N0), e64 = p.malloc<double>(N0), t64 = p.malloc<double>(N0);
  deviceMemory<double> rc64 = p.malloc<double>(N1), ec64 = p.malloc<double>(N1), w1 = p.malloc<double>(N1);
  deviceMemory<float> ec32 = p.malloc<float>(N1), t32 = p.malloc<float>(N0);
  double* tmp = p.o_scalars.ptr() + 20;
  void* ws = p.o_reduceWS.ptr();
  auto A0 = [&](deviceMemory<float>& a, deviceMemory<float>& y) { lev0.Operator(a, y); };

  struct Pair {
    deviceMemory<float> r, eh;
  };
  std::vector<Pair> pairs;
  std::vector<double> nrms;

  auto collect = [&](const std::vector<long>& seeds, bool useNet) {
    for (long seed : seeds) {
      std::vector<deviceMemory<float>> rec;
      L.active = useNet;
      L.fallback = false;
      lev0.record = &rec;
      lev0.recordMax = opt.maxRecordPerSolve;
      el0.BuildRhs(seed, b);
      const int its = fpcgSolve(el0, M0, x, b, w0, 1e-8, 60);
      lev0.record = nullptr;
      L.active = false;  // targets are computed with the Chebyshev cycle
      for (auto& v : rec) {
        LIBP_CALL(libp_convert_f32_to_f64(N0, v.ptr(), r64.ptr(), st));
        LIBP_CALL(libp_norm2sq_f64(N0, r64.ptr(), tmp, ws, st));
        const double nr = std::sqrt(p.readScalars(tmp, 1)[0]);
        if (!(nr > 0) || !std::isfinite(nr)) continue;
        LIBP_CALL(libp_scale_f64(N0, 1.0 / nr, r64.ptr(), st));
        Pair pr;
        pr.r = p.malloc<float>(N0);
        pr.eh = p.malloc<float>(N0);
        LIBP_CALL(libp_convert_f64_to_f32(N0, r64.ptr(), pr.r.ptr(), st));
        // e = A^{-1} r
        fpcgSolve(el0, M0, e64, r64, w0, opt.targetTol, 200);
        // ec = A_1^{-1} R r
        lev0.coarsen(pr.r, o_rhs[1]);
        LIBP_CALL(libp_convert_f32_to_f64(N1, o_rhs[1].ptr(), rc64.ptr(), st));
        fpcgSolve(el1, M1, ec64, rc64, w1, opt.targetTol, 200);
        // eh = e - P ec
        LIBP_CALL(libp_convert_f64_to_f32(N0, e64.ptr(), pr.eh.ptr(), st));
        LIBP_CALL(libp_convert_f64_to_f32(N1, ec64.ptr(), ec32.ptr(), st));
        LIBP_CALL(libp_axpy_f32(N1, 0.f, ec32.ptr(), -1.f, ec32.ptr(), st));
        lev0.prolongate(ec32, pr.eh);
        // nrm = eh.A eh
        A0(pr.eh, t32);
        LIBP_CALL(libp_convert_f32_to_f64(N0, pr.eh.ptr(), t64.ptr(), st));
        LIBP_CALL(libp_convert_f32_to_f64(N0, t32.ptr(), e64.ptr(), st));
        LIBP_CALL(libp_inner_prod_f64(N0, t64.ptr(), e64.ptr(), tmp, ws, st));
        const double nrm = p.readScalars(tmp, 1)[0];
        if (!(nrm > 0) || !std::isfinite(nrm)) continue;
        pairs.push_back(pr);
        nrms.push_back(nrm);
      }
      if (verbose)
        printf("learned smoother: seed %ld (%s smoother) FPCG %d iterations, %zu pairs in total\n", seed,
               useNet ? "learned" : "Chebyshev", its, pairs.size());
    }
  };

  std::vector<double> hist;
  std::mt19937_64 gen(opt.initSeed);
  deviceMemory<double> o_nrm;
  auto train = [&](int steps, float lr0) {
    o_nrm = p.malloc(nrms);
    std::vector<size_t> order(pairs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    L.resetLossSum();
    for (int s = 0; s < steps; ++s) {
      if (s % order.size() == 0) std::shuffle(order.begin(), order.end(), gen);
      const size_t i = order[s % order.size()];
      const float lr = (float)(lr0 * (0.02 + 0.98 * 0.5 * (1.0 + std::cos(M_PI * s / steps))));
      L.trainStep(pairs[i].r.ptr(), pairs[i].eh.ptr(), o_nrm.ptr() + i, lev0.o_invDiagA.ptr(), A0, lr);
      if ((s + 1) % opt.logEvery == 0) {
        hist.push_back(L.stats()[5] / opt.logEvery);
        L.resetLossSum();
        if (verbose) printf("learned smoother: step %d loss %.4f\n", s + 1, hist.back());
      }
    }
  };

  L.initRandom(opt.initSeed);
  collect(opt.seeds, false);
  LIBP_ABORT("learned smoother: no training pairs", pairs.empty());
  train(opt.steps, opt.lr);
  for (int round = 0; round < opt.dagger; ++round) {
    std::vector<long> ds;
    for (long sd : opt.daggerSeeds) ds.push_back(sd + 1000L * round);
    collect(ds, true);
    train(opt.daggerSteps, opt.lr * 0.5f);
  }
  L.active = true;
  L.fallback = false;
  L.status = "trained";
  return hist;
}

/* Least-squares fit of the linear smoother. Samples are white-noise errors v
 * (uniform in [-1,1] at every unmasked node), r = A v, and the error left
 * after an exact coarse correction eh = v - P A_1^{-1} R r (level-1 FPCG with
 * the V-cycle below it to opt.linearCoarseTol). Returns {L2 fit loss
 * |eh - c(r)|^2 / |eh|^2 on a fresh sample}. */
template <>
std::vector<double> multigrid_t<float>::trainLinear(const learnedTrainOptions& opt, int verbose) {
  learnedSmoother_t& L = *levels[0]->learned;
  MGLevel<float>& lev0 = *levels[0];
  elliptic_t& el1 = *ellipticLevels[1];
  platform_t& p = fine.platform;
  void* st = p.s();
  const dlong N0 = ellipticLevels[0]->Ndofs, N1 = el1.Ndofs;
  levelVcycle_t M1(*this, 1);
  deviceMemory<float> v = p.malloc<float>(N0), r = p.malloc<float>(N0), eh = p.malloc<float>(N0);
  deviceMemory<float> ec32 = p.malloc<float>(N1);
  deviceMemory<double> rc64 = p.malloc<double>(N1), ec64 = p.malloc<double>(N1), w1 = p.malloc<double>(N1);
  double* tmp = p.o_scalars.ptr() + 20;
  void* ws = p.o_reduceWS.ptr();
  const bool wasActive = L.active;
  L.active = false;
  // one sample: v from the seed, r = A v, eh = v - P A_1^{-1} R r
  auto sample = [&](unsigned long long seed) {
    LIBP_CALL(libp_random_uniform_f32(N0, seed, v.ptr(), st));
    lev0.Operator(v, r);
    lev0.coarsen(r, o_rhs[1]);
    LIBP_CALL(libp_convert_f32_to_f64(N1, o_rhs[1].ptr(), rc64.ptr(), st));
    const int its = fpcgSolve(el1, M1, ec64, rc64, w1, opt.linearCoarseTol, 200);
    LIBP_CALL(libp_convert_f64_to_f32(N1, ec64.ptr(), ec32.ptr(), st));
    LIBP_CALL(libp_axpy_f32(N1, 0.f, ec32.ptr(), -1.f, ec32.ptr(), st));
    LIBP_HIP(hipMemcpyAsync(eh.ptr(), v.ptr(), N0 * sizeof(float), hipMemcpyDeviceToDevice, p.stream));
    lev0.prolongate(ec32, eh);
    return its;
  };
  std::vector<double> G, B;
  const unsigned long long base = 0x6a09e667f3bcc909ull * (unsigned long long)opt.initSeed;
  for (int k = 0; k < opt.linearSamples; ++k) {
    const int its = sample(base + (unsigned long long)k);
    L.linearGramAccumulate(r.ptr(), eh.ptr(), lev0.o_invDiagA.ptr(), G, B);
    if (verbose) printf("learned smoother (linear): sample %d, level-1 FPCG %d iterations\n", k + 1, its);
  }
  L.linearSolve(G, B, opt.ridge);
  // fit loss on a sample that was not in the fit
  sample(base + (unsigned long long)opt.linearSamples + 7919ull);
  L.correction(r.ptr(), lev0.o_invDiagA.ptr());
  deviceMemory<double> a = p.malloc<double>(N0), b = p.malloc<double>(N0);
  LIBP_CALL(libp_convert_f32_to_f64(N0, eh.ptr(), a.ptr(), st));
  LIBP_CALL(libp_norm2sq_f64(N0, a.ptr(), tmp, ws, st));
  const double ee = p.readScalars(tmp, 1)[0];
  LIBP_CALL(libp_convert_f32_to_f64(N0, L.o_c.ptr(), b.ptr(), st));
  LIBP_CALL(libp_axpy_f64(N0, -1.0, b.ptr(), 1.0, a.ptr(), st));
  LIBP_CALL(libp_norm2sq_f64(N0, a.ptr(), tmp, ws, st));
  const double dd = p.readScalars(tmp, 1)[0];
  const double loss = dd / ee;
  if (verbose) printf("learned smoother (linear): %d samples, ridge %g, validation loss %.4f\n", opt.linearSamples,
                      opt.ridge, loss);
  (void)wasActive;
  L.active = true;
  L.fallback = false;
  L.status = "trained";
  return {loss};
}

template <>
std::vector<double> multigrid_t<double>::trainLinear(const learnedTrainOptions&, int) {
  LIBP_FORCE_ABORT("MULTIGRID SMOOTHER LEARNED needs MULTIGRID PRECISION FP32");
  return {};
}

template <>
std::vector<double> multigrid_t<double>::trainLearned(const learnedTrainOptions&, int) {
  LIBP_FORCE_ABORT("MULTIGRID SMOOTHER LEARNED needs MULTIGRID PRECISION FP32");
  return {};
}

std::vector<double> multigridTrainLearned(MultiGridPrecon& mg, const learnedTrainOptions& opt, int verbose) {
  auto* m = dynamic_cast<multigrid_t<float>*>(mg.impl.get());
  LIBP_ABORT("training needs MULTIGRID PRECISION FP32", !m);
  return m->trainLearned(opt, verbose);
}

learnedSmoother_t* multigridLearned(MultiGridPrecon& mg) {
  auto* m = dynamic_cast<multigrid_t<float>*>(mg.impl.get());
  if (!m || m->levels.empty()) return nullptr;
  return m->levels[0]->learned.get();
}

deviceMemory<float> multigridFineInvDiag32(MultiGridPrecon& mg) {
  auto* m = dynamic_cast<multigrid_t<float>*>(mg.impl.get());
  if (!m || m->levels.empty()) return deviceMemory<float>();
  return m->levels[0]->o_invDiagA;
}

void multigridApplyFineOperator32(MultiGridPrecon& mg, deviceMemory<float>& x, deviceMemory<float>& y) {
  auto* m = dynamic_cast<multigrid_t<float>*>(mg.impl.get());
  LIBP_ABORT("needs an FP32 p-MG preconditioner", !m || m->levels.empty());
  m->levels[0]->Operator(x, y);
}

namespace {
template <typename T>
void levelOp(multigrid_t<T>& m, int l, int op, const void* in, void* out, int flag) {
  LIBP_ABORT("no such p-MG level", l < 0 || l + 1 >= (int)m.ellipticLevels.size());
  MGLevel<T>& L = *m.levels[l];
  const dlong Nf = m.ellipticLevels[l]->Ndofs, Nc = m.ellipticLevels[l + 1]->Ndofs;
  T* pin = const_cast<T*>(static_cast<const T*>(in));
  T* pout = static_cast<T*>(out);
  deviceMemory<T> x = deviceMemory<T>::wrap(pin, op == 2 ? Nc : Nf);
  deviceMemory<T> y = deviceMemory<T>::wrap(pout, op == 1 ? Nc : Nf);
  switch (op) {
    case 0: L.Operator(x, y); break;
    case 1: L.coarsen(x, y); break;
    case 2: L.prolongate(x, y); break;
    case 3: L.smooth(x, y, flag != 0); break;
    case 4:
      LIBP_HIP(hipMemcpyAsync(pout, L.o_invDiagA.ptr(), Nf * sizeof(T), hipMemcpyDeviceToDevice,
                              m.fine.platform.stream));
      break;
    default: LIBP_FORCE_ABORT("unknown p-MG level operation");
  }
}
}  // namespace

int multigridPrecisionBits(MultiGridPrecon& mg) {
  if (dynamic_cast<multigrid_t<float>*>(mg.impl.get())) return 32;
  if (dynamic_cast<multigrid_t<double>*>(mg.impl.get())) return 64;
  return 0;
}

long multigridLevelNdofs(MultiGridPrecon& mg, int l) {
  std::vector<std::shared_ptr<elliptic_t>>* lv = nullptr;
  if (auto* m = dynamic_cast<multigrid_t<float>*>(mg.impl.get())) lv = &m->ellipticLevels;
  if (auto* m = dynamic_cast<multigrid_t<double>*>(mg.impl.get())) lv = &m->ellipticLevels;
  if (!lv || l < 0 || l >= (int)lv->size()) return -1;
  return (*lv)[l]->Ndofs;
}

void multigridLevelOp(MultiGridPrecon& mg, int l, int op, const void* in, void* out, int flag) {
  if (auto* m = dynamic_cast<multigrid_t<float>*>(mg.impl.get())) return levelOp(*m, l, op, in, out, flag);
  if (auto* m = dynamic_cast<multigrid_t<double>*>(mg.impl.get())) return levelOp(*m, l, op, in, out, flag);
  LIBP_FORCE_ABORT("no p-MG preconditioner");
}

void multigridLevelLambda(MultiGridPrecon& mg, int l, double* lambda) {
  auto get = [&](auto* m) {
    LIBP_ABORT("no such p-MG level", l < 0 || l >= (int)m->levels.size());
    lambda[0] = m->levels[l]->lambda1;
    lambda[1] = m->levels[l]->lambda0;
  };
  if (auto* m = dynamic_cast<multigrid_t<float>*>(mg.impl.get())) return get(m);
  if (auto* m = dynamic_cast<multigrid_t<double>*>(mg.impl.get())) return get(m);
  LIBP_FORCE_ABORT("no p-MG preconditioner");
}

amgSolver_t* multigridBaseAmg(MultiGridPrecon& mg) {
  if (auto* m = dynamic_cast<multigrid_t<float>*>(mg.impl.get())) return m->coarse->amg.get();
  if (auto* m = dynamic_cast<multigrid_t<double>*>(mg.impl.get())) return m->coarse->amg.get();
  return nullptr;
}

int multigridCoarseIterations(MultiGridPrecon& mg) {
  if (auto* m = dynamic_cast<multigrid_t<float>*>(mg.impl.get())) return m->coarse->iterations();
  if (auto* m = dynamic_cast<multigrid_t<double>*>(mg.impl.get())) return m->coarse->iterations();
  return -1;
}

}  // namespace libp
