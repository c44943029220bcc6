// CPU oracle (test infrastructure only): BLAS-1 reductions with the reference's
// block structure, PCG/FPCG, the p-multigrid V-cycle, and the end-to-end run.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

std::vector<std::complex<double>> hessenbergEigs(int n, const std::vector<double>& Hcm);

/* Two-stage block reduction of libs/linAlg/okl/linAlgInnerProd.okl:28-100
 * with blocksize 256 and Nblock = min(ceil(N/256), 256) (linAlg.cpp:192-207,
 * include/linAlg.hpp:179), evaluated in OCCA Serial order. */
template <typename F>
static double okl_reduce(dlong N, F f) {
  const int bs = 256;
  int Nblock = (N + bs - 1) / bs;
  Nblock = (Nblock > bs) ? bs : Nblock;
  if (Nblock == 0) return 0.0;
  std::vector<double> dot(Nblock), s(bs);
  for (int b = 0; b < Nblock; ++b) {
    for (int t = 0; t < bs; ++t) {
      dlong id = t + b * bs;
      double r = 0.0;
      while (id < N) {
        r += f(id);
        id += bs * Nblock;
      }
      s[t] = r;
    }
    for (int st = 128; st >= 2; st /= 2)
      for (int t = 0; t < st; ++t) s[t] += s[t + st];
    dot[b] = s[0] + s[1];
  }
  for (int t = 0; t < bs; ++t) {
    dlong id = t;
    double r = 0.0;
    while (id < Nblock) {
      r += dot[id];
      id += bs;
    }
    s[t] = r;
  }
  for (int st = 128; st >= 2; st /= 2)
    for (int t = 0; t < st; ++t) s[t] += s[t + st];
  return s[0] + s[1];
}

double linalg_innerProd(dlong N, const double* x, const double* y) {
  return okl_reduce(N, [&](dlong i) { return x[i] * y[i]; });
}
double linalg_norm2(dlong N, const double* x) {
  return std::sqrt(okl_reduce(N, [&](dlong i) { return x[i] * x[i]; }));
}
double linalg_sum(dlong N, const double* x) {
  return okl_reduce(N, [&](dlong i) { return x[i]; });
}

/* linAlgAXPY.okl:27-51 and linAlgAMXPY.okl:27-63. */
static void axpy(dlong N, double alpha, const double* x, double beta, double* y) {
  for (dlong n = 0; n < N; ++n) y[n] = (beta != 0) ? alpha * x[n] + beta * y[n] : alpha * x[n];
}
static void amxpy(dlong N, double alpha, const double* a, const double* x, double beta, double* y) {
  for (dlong n = 0; n < N; ++n) y[n] = (beta != 0) ? alpha * a[n] * x[n] + beta * y[n] : alpha * a[n] * x[n];
}
static void amx(dlong N, double alpha, const double* a, double* x) {
  for (dlong n = 0; n < N; ++n) x[n] = alpha * a[n] * x[n];
}

/* updatePCG (libs/linearSolver/okl/linearSolverUpdatePCG.okl:29-75) with
 * blocksize 512, Nblocks = min(ceil(N/512), 512), partials summed on the host
 * in block order (linearSolverPCG.cpp:153-171). */
static double updatePCG(dlong N, const double* p, const double* Ap, double alpha, double* x, double* r) {
  const int bs = 512;
  int Nblocks = (N + bs - 1) / bs;
  Nblocks = std::min(Nblocks, bs);
  std::vector<double> s(bs);
  double total = 0;
  for (int b = 0; b < Nblocks; ++b) {
    for (int t = 0; t < bs; ++t) {
      dlong id = t + b * bs;
      s[t] = 0.0;
      while (id < N) {
        double rn = r[id];
        x[id] += alpha * p[id];
        rn -= alpha * Ap[id];
        s[t] += rn * rn;
        r[id] = rn;
        id += bs * Nblocks;
      }
    }
    for (int st = 256; st >= 2; st /= 2)
      for (int t = 0; t < st; ++t) s[t] += s[t + st];
    total += s[0] + s[1];
  }
  return total;
}

/* libs/linearSolver/linearSolverPCG.cpp:67-151. pureRelative replaces the
 * reference's max(tol^2*r0^2, tol^2) floor with tol^2*r0^2 (no reference
 * counterpart; used for time-to-relative-residual measurements). */
PCGResult pcg_solve(dlong N, OpFn& A, OpFn& M, double* x, double* r, double tol, int maxit, bool flexible,
                    bool stopRhsNorm, bool pureRelative) {
  PCGResult res;
  std::vector<double> p(N, 0.0), z(N, 0.0), Ax(N, 0.0), Ap(N, 0.0);
  double rdotz1 = 0, rdotz2 = 0, alpha = 0, beta = 0, pAp = 0, rdotr0 = 0, TOL = 0;
  if (stopRhsNorm) {
    double normb = linalg_norm2(N, r);
    TOL = std::max(tol * tol * normb * normb, tol * tol);
  }
  A.apply(x, Ax.data());
  axpy(N, -1.0, Ax.data(), 1.0, r);
  rdotr0 = linalg_norm2(N, r);
  rdotr0 = rdotr0 * rdotr0;
  if (!stopRhsNorm) TOL = pureRelative ? tol * tol * rdotr0 : std::max(tol * tol * rdotr0, tol * tol);
  res.rdotr0 = rdotr0;
  res.resHistory.push_back(std::sqrt(rdotr0));
  int iter;
  for (iter = 0; iter < maxit; ++iter) {
    if (((iter == 0) && (rdotr0 == 0.0)) || ((iter > 0) && (rdotr0 <= TOL))) break;
    M.apply(r, z.data());
    rdotz2 = rdotz1;
    rdotz1 = linalg_innerProd(N, r, z.data());
    if (flexible) {
      double zdotAp = linalg_innerProd(N, z.data(), Ap.data());
      beta = (iter == 0) ? 0.0 : -alpha * zdotAp / rdotz2;
    } else {
      beta = (iter == 0) ? 0.0 : rdotz1 / rdotz2;
    }
    axpy(N, 1.0, z.data(), beta, p.data());
    A.apply(p.data(), Ap.data());
    pAp = linalg_innerProd(N, p.data(), Ap.data());
    alpha = rdotz1 / pAp;
    rdotr0 = updatePCG(N, p.data(), Ap.data(), alpha, x, r);
    res.resHistory.push_back(std::sqrt(std::max(rdotr0, 0.0)));
  }
  res.iters = iter;
  res.rdotrFinal = rdotr0;
  return res;
}

/* Fused multi-dot of the non-blocking solvers: blocksize 512, Nblocks =
 * min(ceil(N/512), 512), thread t of block b sums n = t + 512 b step
 * 512 Nblocks, tree to two values, partials summed on the host in block
 * order (linearSolverUpdateNBPCG.okl / linearSolverUpdateNBFPCG.okl,
 * linearSolverNBPCG.cpp:170-224, linearSolverNBFPCG.cpp:183-239). f(n, out)
 * performs the pointwise update of entry n and adds its K products to out. */
template <int K, typename F>
static void nbDots(dlong N, F f, double* dots) {
  const int bs = 512;
  int Nblocks = (N + bs - 1) / bs;
  Nblocks = std::min(Nblocks, bs);
  for (int k = 0; k < K; ++k) dots[k] = 0.0;
  std::vector<double> s((size_t)K * bs);
  for (int b = 0; b < Nblocks; ++b) {
    for (int t = 0; t < bs; ++t) {
      double acc[K];
      for (int k = 0; k < K; ++k) acc[k] = 0.0;
      for (dlong id = t + (dlong)b * bs; id < N; id += (dlong)bs * Nblocks) f(id, acc);
      for (int k = 0; k < K; ++k) s[(size_t)k * bs + t] = acc[k];
    }
    for (int st = 256; st >= 2; st /= 2)
      for (int t = 0; t < st; ++t)
        for (int k = 0; k < K; ++k) s[(size_t)k * bs + t] += s[(size_t)k * bs + t + st];
    for (int k = 0; k < K; ++k) dots[k] += s[(size_t)k * bs] + s[(size_t)k * bs + 1];
  }
}

/* LinearSolver::nbpcg::Solve (libs/linearSolver/linearSolverNBPCG.cpp:71-160). */
PCGResult nbpcg_solve(dlong N, OpFn& A, OpFn& M, double* x, double* r, double tol, int maxit, bool pureRelative) {
  PCGResult res;
  std::vector<double> p(N, 0.0), s(N, 0.0), z(N, 0.0), Z(N, 0.0), S(N, 0.0), Ax(N, 0.0);
  double dots[3];
  A.apply(x, Ax.data());
  axpy(N, -1.0, Ax.data(), 1.0, r);
  M.apply(r, z.data());
  double alpha = 0;
  auto update2 = [&](double a) {
    nbDots<3>(N, [&](dlong n, double* o) {
      double rn = r[n] - a * s[n], zn = z[n] - a * S[n];
      o[0] += rn * zn;
      o[1] += zn * zn;
      o[2] += rn * rn;
      r[n] = rn;
      z[n] = zn;
    }, dots);
  };
  update2(alpha);
  A.apply(z.data(), Z.data());
  double gamma = dots[0], rdotr = dots[2], gamma1 = 0, beta = 0, delta = 0;
  const double TOL = pureRelative ? tol * tol * rdotr : std::max(tol * tol * rdotr, tol * tol);
  res.rdotr0 = rdotr;
  res.resHistory.push_back(std::sqrt(rdotr));
  int iter;
  for (iter = 0; iter < maxit; ++iter) {
    if (rdotr <= TOL) break;
    nbDots<1>(N, [&](dlong n, double* o) {
      const double pn = z[n] + beta * p[n], sn = Z[n] + beta * s[n];
      o[0] += pn * sn;
      p[n] = pn;
      s[n] = sn;
    }, dots);
    M.apply(s.data(), S.data());
    delta = dots[0];
    alpha = gamma / delta;
    update2(alpha);
    axpy(N, alpha, p.data(), 1.0, x);
    A.apply(z.data(), Z.data());
    gamma1 = gamma;
    gamma = dots[0];
    rdotr = dots[2];
    beta = gamma / gamma1;
    res.resHistory.push_back(std::sqrt(std::max(rdotr, 0.0)));
  }
  res.iters = iter;
  res.rdotrFinal = rdotr;
  return res;
}

/* LinearSolver::nbfpcg::Solve (libs/linearSolver/linearSolverNBFPCG.cpp:71-181), Sanan's pipelined flexible CG. */
PCGResult nbfpcg_solve(dlong N, OpFn& A, OpFn& M, double* x, double* r, double tol, int maxit, bool pureRelative) {
  PCGResult res;
  std::vector<double> u(N), p(N), w(N), m(N), n_(N), s(N), q(N), z(N), Ax(N);
  double dots[4];
  A.apply(x, Ax.data());
  axpy(N, -1.0, Ax.data(), 1.0, r);
  M.apply(r, u.data());
  p = u;
  A.apply(p.data(), w.data());
  nbDots<3>(N, [&](dlong i, double* o) {
    o[0] += u[i] * r[i];
    o[1] += u[i] * w[i];
    o[2] += r[i] * r[i];
  }, dots);
  M.apply(w.data(), m.data());
  A.apply(m.data(), n_.data());
  s = w;
  q = m;
  z = n_;
  double gamma = dots[0], delta = dots[1], rdotr = dots[2], eta = delta, alpha = gamma / eta, beta = 0;
  const double TOL = pureRelative ? tol * tol * rdotr : std::max(tol * tol * rdotr, tol * tol);
  res.rdotr0 = rdotr;
  res.resHistory.push_back(std::sqrt(rdotr));
  int iter;
  for (iter = 0; iter < maxit; ++iter) {
    if (rdotr <= TOL) break;
    nbDots<4>(N, [&](dlong i, double* o) {
      const double xn = x[i] + alpha * p[i], rn = r[i] - alpha * s[i];
      const double un = u[i] - alpha * q[i], wn = w[i] - alpha * z[i];
      o[0] += un * rn;
      o[1] += un * s[i];
      o[2] += un * wn;
      o[3] += rn * rn;
      x[i] = xn;
      r[i] = rn;
      u[i] = un;
      w[i] = wn;
    }, dots);
    for (dlong i = 0; i < N; ++i) n_[i] = 1.0 * w[i] - 1.0 * r[i];  // zaxpy
    M.apply(n_.data(), m.data());
    axpy(N, 1.0, u.data(), 1.0, m.data());
    A.apply(m.data(), n_.data());
    gamma = dots[0];
    beta = -dots[1] / eta;
    delta = dots[2];
    rdotr = dots[3];
    axpy(N, 1.0, u.data(), beta, p.data());
    axpy(N, 1.0, w.data(), beta, s.data());
    axpy(N, 1.0, m.data(), beta, q.data());
    axpy(N, 1.0, n_.data(), beta, z.data());
    eta = delta - beta * beta * eta;
    alpha = gamma / eta;
    res.resHistory.push_back(std::sqrt(std::max(rdotr, 0.0)));
  }
  res.iters = iter;
  res.rdotrFinal = rdotr;
  return res;
}

/* ----------------------------- p-multigrid ----------------------------- */
namespace {
struct EllOp : OpFn {
  const Elliptic* el;
  std::vector<double> scratch;
  explicit EllOp(const Elliptic* e) : el(e) {}
  void apply(const double* x, double* y) override { el->Operator(x, y, scratch); }
};
struct DiagOp : OpFn {
  const std::vector<double>* inv;
  explicit DiagOp(const std::vector<double>* i) : inv(i) {}
  void apply(const double* x, double* y) override { amxpy((dlong)inv->size(), 1.0, inv->data(), x, 0.0, y); }
};
struct IdOp : OpFn {
  dlong N;
  explicit IdOp(dlong n) : N(n) {}
  void apply(const double* x, double* y) override { std::copy(x, x + N, y); }
};
}  // namespace

/* MGLevel::maxEigSmoothAx (ellipticPreconMultiGridLevel.cpp:393-473). */
double maxEigSmoothAx(const Elliptic& el, const std::vector<double>& invDiagA, Drand48& drng) {
  const dlong N = el.Ndofs;
  int k = 10;
  if (k > N) k = N;
  std::vector<double> H(k * k, 0.0);
  std::vector<std::vector<double>> V(k + 1, std::vector<double>(N, 0.0));
  std::vector<double> Vx(N), AVx(N), scratch;
  for (dlong i = 0; i < N; i++) Vx[i] = drng.next();
  double norm_vo = linalg_norm2(N, Vx.data());
  axpy(N, 1. / norm_vo, Vx.data(), 0.0, V[0].data());
  for (int j = 0; j < k; j++) {
    el.Operator(V[j].data(), AVx.data(), scratch);
    amxpy(N, 1.0, invDiagA.data(), AVx.data(), 0.0, V[j + 1].data());
    for (int i = 0; i <= j; i++) {
      double hij = linalg_innerProd(N, V[i].data(), V[j + 1].data());
      axpy(N, -hij, V[i].data(), 1.0, V[j + 1].data());
      H[i + j * k] = hij;
    }
    if (j + 1 < k) {
      double norm_vj = linalg_norm2(N, V[j + 1].data());
      for (dlong n = 0; n < N; ++n) V[j + 1][n] = (1.0 / norm_vj) * V[j + 1][n];
      H[j + 1 + j * k] = norm_vj;
    }
  }
  auto ev = hessenbergEigs(k, H);
  double rho = 0.;
  for (auto& l : ev) rho = std::max(rho, std::abs(l));
  return rho;
}

/* Degree schedule of MultiGridPrecon (ellipticPreconMultiGrid.cpp:45-131),
 * extended with a "CUSTOM" option reading [MULTIGRID DEGREES] (e.g. 7,3,1),
 * which the reference cannot express. */
static std::vector<int> mgDegrees(const Settings& s, int N) {
  std::vector<int> deg;
  if (s.compare("MULTIGRID COARSENING", "CUSTOM")) {
    std::string v = s.get("MULTIGRID DEGREES");
    size_t p = 0;
    while (p < v.size()) {
      size_t q = v.find(',', p);
      if (q == std::string::npos) q = v.size();
      deg.push_back(std::stoi(v.substr(p, q - p)));
      p = q + 1;
    }
    if (deg.empty() || deg[0] != N || deg.back() != 1) throw std::runtime_error("bad MULTIGRID DEGREES");
    return deg;
  }
  int Nf = N, Nc = N, NpFine = (N + 1) * (N + 1) * (N + 1), NpCoarse = NpFine;
  deg.push_back(N);
  while (Nc > 1) {
    if (s.compare("MULTIGRID COARSENING", "ALLDEGREES")) {
      Nc = Nf - 1;
    } else if (s.compare("MULTIGRID COARSENING", "HALFDEGREES")) {
      Nc = std::max(1, (Nf + 1) / 2);
    } else {
      while (NpCoarse > NpFine / 2 && Nc > 1) {
        Nc--;
        NpCoarse = (Nc + 1) * (Nc + 1) * (Nc + 1);
      }
    }
    NpCoarse = (Nc + 1) * (Nc + 1) * (Nc + 1);
    deg.push_back(Nc);
    Nf = Nc;
    NpFine = NpCoarse;
  }
  return deg;
}

void MultiGrid::Setup(Elliptic& fine, GlibcRand& rng, Drand48& drng, const std::vector<int>& degrees) {
  // level 0 is the fine operator itself (SetupNewDegree returns *this)
  levels.clear();
  levels.emplace_back(new Elliptic(fine));
  drng.seed(0);  // parAlmondKernels.cpp:55-58 (rank 0)
  const Settings& s = fine.settings;
  for (size_t l = 1; l < degrees.size(); ++l) {
    // The level's operator is built before the previous level's smoother
    // setup only in the sense of data sharing; in the reference the MGLevel
    // for degree degrees[l-1] is constructed after its elliptic and before
    // the next degree's elliptic, so rand() and drand48() interleave as:
    //   ogs(deg l-1) [fine: already done], Arnoldi(deg l-1), ogs(deg l), ...
    MGLevelO lev;
    lev.el = levels[l - 1].get();
    lev.Nc = degrees[l];
    DegreeRaiseMatrix1D(degrees[l], degrees[l - 1], lev.P);
    std::vector<double> diagA;
    lev.el->BuildOperatorDiagonal(diagA);
    lev.invDiagA.resize(diagA.size());
    for (size_t n = 0; n < diagA.size(); ++n) lev.invDiagA[n] = 1.0 / diagA[n];
    if (s.compare("MULTIGRID SMOOTHER", "CHEBYSHEV")) {
      lev.stype = 1;
      lev.ChebyshevIterations = s.has("MULTIGRID CHEBYSHEV DEGREE") ? s.geti("MULTIGRID CHEBYSHEV DEGREE") : 2;
      double rho = maxEigSmoothAx(*lev.el, lev.invDiagA, drng);
      lev.lambda1 = rho;
      lev.lambda0 = rho / 10.;
    } else {
      lev.stype = 0;
      double rho = maxEigSmoothAx(*lev.el, lev.invDiagA, drng);
      lev.lambda0 = (4. / 3.) / rho;
      for (auto& v : lev.invDiagA) v *= lev.lambda0;
    }
    mg.push_back(lev);
    // next (coarser) elliptic
    Elliptic* elC = new Elliptic();
    elC->Setup(fine.mesh.SetupNewDegree(degrees[l]), s, fine.lambda, rng);
    levels.emplace_back(elC);
    mg.back().elC = elC;
  }
  // base level: the reference's parAlmond AMG on the assembled N=1 operator
  // (ellipticPreconMultiGrid.cpp:119-150), continuing the drand48 stream
  if (s.has("MULTIGRID COARSE SOLVER") && s.compare("MULTIGRID COARSE SOLVER", "AMG")) {
    const Elliptic& C = *levels.back();
    amg.reset(new AmgO());
    amg->Setup(BuildOperatorMatrixContinuousHex3D(C), s, drng, C.allNeumann != 0, C.allNeumannPenalty);
  }
  // coarse Jacobi-PCG data
  std::vector<double> diagC;
  levels.back()->BuildOperatorDiagonal(diagC);
  invDiagCoarse.assign(1, std::vector<double>(diagC.size()));
  for (size_t n = 0; n < diagC.size(); ++n) invDiagCoarse[0][n] = 1.0 / diagC[n];
  rhsL.resize(levels.size());
  xL.resize(levels.size());
  for (size_t l = 0; l < levels.size(); ++l) {
    rhsL[l].assign(levels[l]->Ndofs, 0.0);
    xL[l].assign(levels[l]->Ndofs, 0.0);
  }
}

void smoothChebyshev(const MGLevelO& L, const double* r, double* X, bool xIsZero) {
  const Elliptic& el = *L.el;
  const dlong N = el.Ndofs;
  const double theta = 0.5 * (L.lambda1 + L.lambda0);
  const double delta = 0.5 * (L.lambda1 - L.lambda0);
  const double invTheta = 1.0 / theta;
  const double sigma = theta / delta;
  double rho_n = 1. / sigma, rho_np1;
  std::vector<double> RES(N), Ad(N), d(N), scratch;
  if (xIsZero) {
    amxpy(N, 1.0, L.invDiagA.data(), r, 0.0, RES.data());
    axpy(N, invTheta, RES.data(), 0.0, d.data());
  } else {
    el.Operator(X, RES.data(), scratch);
    axpy(N, 1.0, r, -1.0, RES.data());
    amx(N, 1.0, L.invDiagA.data(), RES.data());
    axpy(N, invTheta, RES.data(), 0.0, d.data());
  }
  for (int k = 0; k < L.ChebyshevIterations; k++) {
    if (xIsZero && (k == 0))
      axpy(N, 1.0, d.data(), 0.0, X);
    else
      axpy(N, 1.0, d.data(), 1.0, X);
    el.Operator(d.data(), Ad.data(), scratch);
    amxpy(N, -1.0, L.invDiagA.data(), Ad.data(), 1.0, RES.data());
    rho_np1 = 1.0 / (2. * sigma - rho_n);
    double rhoDivDelta = 2.0 * rho_np1 / delta;
    axpy(N, rhoDivDelta, RES.data(), rho_np1 * rho_n, d.data());
    rho_n = rho_np1;
  }
  axpy(N, 1.0, d.data(), 1.0, X);
}

void smoothJacobi(const MGLevelO& L, const double* r, double* X, bool xIsZero) {
  const Elliptic& el = *L.el;
  const dlong N = el.Ndofs;
  if (xIsZero) {
    amxpy(N, 1.0, L.invDiagA.data(), r, 0.0, X);
    return;
  }
  std::vector<double> RES(N), scratch;
  el.Operator(X, RES.data(), scratch);
  axpy(N, 1.0, r, -1.0, RES.data());
  amxpy(N, 1.0, L.invDiagA.data(), RES.data(), 1.0, X);
}

/* ellipticPartialPreconCoarsenHex3D (okl/ellipticPreconCoarsenHex3D.okl:209-294)
 * applied through MGLevel::coarsen (ellipticPreconMultiGridLevel.cpp:41-82). */
void coarsen(const MGLevelO& L, const double* X, double* Rx) {
  const Elliptic& F = *L.el;
  const Elliptic& C = *L.elC;
  const int NqF = F.mesh.Nq, NqC = L.Nc + 1, NpF = F.mesh.Np, NpC = NqC * NqC * NqC;
  std::vector<double> wx(F.Ndofs);
  amxpy(F.Ndofs, 1.0, F.weightG.data(), X, 0.0, wx.data());
  std::vector<double> RxL((size_t)F.mesh.Nelements * NpC);
  std::vector<double> qfff(NpF), qcff(NqC * NqF * NqF), qccf(NqC * NqC * NqF);
  const double* P = L.P.data();  // s_RT[j][i] = RT[j*NqC+i]
  for (dlong e = 0; e < F.mesh.Nelements; ++e) {
    for (int n = 0; n < NpF; ++n) {
      dlong id = F.GlobalToLocal[(size_t)e * NpF + n];
      qfff[n] = (id != -1) ? wx[id] : 0.0;
    }
    for (int k = 0; k < NqC; ++k)
      for (int j = 0; j < NqF; ++j)
        for (int i = 0; i < NqF; ++i) {
          double res = 0;
          for (int m = 0; m < NqF; ++m) res += P[m * NqC + k] * qfff[m * NqF * NqF + j * NqF + i];
          qcff[k * NqF * NqF + j * NqF + i] = res;
        }
    for (int k = 0; k < NqC; ++k)
      for (int j = 0; j < NqC; ++j)
        for (int i = 0; i < NqF; ++i) {
          double res = 0;
          for (int m = 0; m < NqF; ++m) res += P[m * NqC + j] * qcff[k * NqF * NqF + m * NqF + i];
          qccf[k * NqC * NqF + j * NqF + i] = res;
        }
    for (int k = 0; k < NqC; ++k)
      for (int j = 0; j < NqC; ++j)
        for (int i = 0; i < NqC; ++i) {
          double r = 0;
          for (int m = 0; m < NqF; ++m) r += P[m * NqC + i] * qccf[k * NqC * NqF + j * NqF + m];
          RxL[(size_t)e * NpC + i + j * NqC + k * NqC * NqC] = r;
        }
  }
  C.ogsMasked.gatherLocal.GatherAdd(Rx, RxL.data(), true);
}

/* ellipticPartialPreconProlongateHex3D (okl/ellipticPreconProlongateHex3D.okl:216-301)
 * applied through MGLevel::prolongate (ellipticPreconMultiGridLevel.cpp:84-125). */
void prolongate(const MGLevelO& L, const double* Xc, double* Px) {
  const Elliptic& F = *L.el;
  const Elliptic& C = *L.elC;
  const int NqF = F.mesh.Nq, NqC = L.Nc + 1, NpF = F.mesh.Np, NpC = NqC * NqC * NqC;
  const double* P = L.P.data();  // s_P[j][i] = P[i + j*NqC]
  std::vector<double> PxL((size_t)F.mesh.Nelements * NpF), PxG(F.Ndofs);
  std::vector<double> qccc(NpC), qccf(NqC * NqC * NqF), qcff(NqC * NqF * NqF);
  for (dlong e = 0; e < C.mesh.Nelements; ++e) {
    for (int n = 0; n < NpC; ++n) {
      dlong id = C.GlobalToLocal[(size_t)e * NpC + n];
      qccc[n] = (id != -1) ? Xc[id] : 0.0;
    }
    for (int k = 0; k < NqC; ++k)
      for (int j = 0; j < NqC; ++j)
        for (int i = 0; i < NqF; ++i) {
          double res = 0;
          for (int m = 0; m < NqC; ++m) res += P[i * NqC + m] * qccc[k * NqC * NqC + j * NqC + m];
          qccf[k * NqC * NqF + j * NqF + i] = res;
        }
    for (int k = 0; k < NqC; ++k)
      for (int j = 0; j < NqF; ++j)
        for (int i = 0; i < NqF; ++i) {
          double res = 0;
          for (int m = 0; m < NqC; ++m) res += P[j * NqC + m] * qccf[k * NqC * NqF + m * NqF + i];
          qcff[k * NqF * NqF + j * NqF + i] = res;
        }
    for (int k = 0; k < NqF; ++k)
      for (int j = 0; j < NqF; ++j)
        for (int i = 0; i < NqF; ++i) {
          double res = 0;
          for (int m = 0; m < NqC; ++m) res += P[k * NqC + m] * qcff[m * NqF * NqF + j * NqF + i];
          PxL[(size_t)e * NpF + i + j * NqF + k * NqF * NqF] = res;
        }
  }
  F.ogsMasked.gatherLocal.GatherAdd(PxG.data(), PxL.data(), false);
  axpy(F.Ndofs, 1.0, PxG.data(), 1.0, Px);
}

/* multigrid_t::vcycle (libs/parAlmond/parAlmondVcycle.cpp:34-60). The base
 * level is Jacobi-PCG on the N=1 operator, from a zero initial guess, to a
 * pure relative tolerance coarseTol. */
void MultiGrid::vcycle(int k, std::vector<double>& rhs, std::vector<double>& x) {
  const int base = (int)levels.size() - 1;
  if (k == base && amg) {
    amg->vcycle(0, rhs.data(), x.data());
    return;
  }
  if (k == base) {
    Elliptic& C = *levels[base];
    EllOp A(&C);
    DiagOp M(&invDiagCoarse[0]);
    std::fill(x.begin(), x.end(), 0.0);
    std::vector<double> r = rhs;
    pcg_solve(C.Ndofs, A, M, x.data(), r.data(), coarseTol, coarseMaxIt, false, false, true);
    return;
  }
  const MGLevelO& L = mg[k];
  const Elliptic& el = *L.el;
  std::vector<double> RES(el.Ndofs), scratch;
  if (L.stype == 1) smoothChebyshev(L, rhs.data(), x.data(), true);
  else smoothJacobi(L, rhs.data(), x.data(), true);
  el.Operator(x.data(), RES.data(), scratch);
  axpy(el.Ndofs, 1.0, rhs.data(), -1.0, RES.data());
  coarsen(L, RES.data(), rhsL[k + 1].data());
  vcycle(k + 1, rhsL[k + 1], xL[k + 1]);
  prolongate(L, xL[k + 1].data(), x.data());
  if (L.stype == 1) smoothChebyshev(L, rhs.data(), x.data(), false);
  else smoothJacobi(L, rhs.data(), x.data(), false);
}

void MultiGrid::apply(const double* r, double* z) {
  std::vector<double> rr(r, r + levels[0]->Ndofs), zz(levels[0]->Ndofs, 0.0);
  vcycle(0, rr, zz);
  std::copy(zz.begin(), zz.end(), z);
}

/* elliptic_t::Run (solvers/elliptic/src/ellipticRun.cpp:30-248) for the C0
 * hex path with NONE / JACOBI / MULTIGRID preconditioning. */
RunResult elliptic_run(const Settings& s, int verbose) {
  RunResult out;
  GlibcRand rng;
  Drand48 drng;
  Mesh mesh;
  mesh.Setup(s);
  Elliptic el;
  el.Setup(mesh, s, s.getd("LAMBDA"), rng);
  out.Ngather = el.Ndofs;
  const dlong N = el.Ndofs;
  const size_t NL = (size_t)mesh.Nelements * mesh.Np;

  std::vector<double> rL, xL(NL, 0.0), r(N), x(N);
  el.Rhs(rL);
  el.ogsMasked.gatherLocal.GatherAdd(r.data(), rL.data(), true);
  el.ogsMasked.gatherLocal.GatherAdd(x.data(), xL.data(), false);
  if (el.allNeumann) el.ZeroMean(r.data());

  EllOp A(&el);
  std::unique_ptr<OpFn> M;
  std::vector<double> invDiag;
  MultiGrid mgp;
  struct MGOp : OpFn {
    MultiGrid* mg;
    const Elliptic* el;
    void apply(const double* x, double* y) override {
      mg->apply(x, y);
      if (el->allNeumann) el->ZeroMean(y);
    }
  } mgop;
  struct JacOp : OpFn {
    const std::vector<double>* inv;
    const Elliptic* el;
    void apply(const double* x, double* y) override {
      amxpy((dlong)inv->size(), 1.0, inv->data(), x, 0.0, y);
      if (el->allNeumann) el->ZeroMean(y);
    }
  } jac;
  IdOp ident(N);
  OpFn* Mp = &ident;
  if (s.compare("PRECONDITIONER", "JACOBI")) {
    std::vector<double> diagA;
    el.BuildOperatorDiagonal(diagA);
    invDiag.resize(N);
    for (dlong n = 0; n < N; ++n) invDiag[n] = 1.0 / diagA[n];
    jac.inv = &invDiag;
    jac.el = &el;
    Mp = &jac;
  } else if (s.compare("PRECONDITIONER", "MULTIGRID")) {
    if (s.has("MULTIGRID COARSE TOLERANCE")) mgp.coarseTol = s.getd("MULTIGRID COARSE TOLERANCE");
    if (s.has("MULTIGRID COARSE MAXIT")) mgp.coarseMaxIt = s.geti("MULTIGRID COARSE MAXIT");
    mgp.Setup(el, rng, drng, mgDegrees(s, mesh.N));
    mgop.mg = &mgp;
    mgop.el = &el;
    Mp = &mgop;
  }
  const bool flexible = s.compare("LINEAR SOLVER", "FPCG");
  const bool rhsNorm = s.compare("LINEAR SOLVER STOPPING CRITERION", "ABS/REL-RHS-2NORM");
  const bool pureRel = s.compare("LINEAR SOLVER STOPPING CRITERION", "RELATIVE") &&
                       !s.compare("LINEAR SOLVER STOPPING CRITERION", "ABS/REL");
  const int maxit = s.has("MAXIMUM ITERATIONS") ? s.geti("MAXIMUM ITERATIONS") : 5000;
  const double tol = s.has("LINEAR SOLVER TOLERANCE") ? s.getd("LINEAR SOLVER TOLERANCE") : 1.0e-8;
  PCGResult pr;
  if (s.compare("LINEAR SOLVER", "NBFPCG"))
    pr = nbfpcg_solve(N, A, *Mp, x.data(), r.data(), tol, maxit, pureRel);
  else if (s.compare("LINEAR SOLVER", "NBPCG"))
    pr = nbpcg_solve(N, A, *Mp, x.data(), r.data(), tol, maxit, pureRel);
  else
    pr = pcg_solve(N, A, *Mp, x.data(), r.data(), tol, maxit, flexible, rhsNorm, pureRel);
  out.iters = pr.iters;
  if (verbose) printf("oracle: %d iterations, final |r| %.6e\n", pr.iters, std::sqrt(std::max(pr.rdotrFinal, 0.0)));
  el.ogsMasked.gatherLocal.Scatter(xL.data(), x.data(), false);
  el.AddBC(xL);
  std::vector<double> MxL(NL);
  for (size_t n = 0; n < NL; ++n) MxL[n] = mesh.wJ[n] * xL[n];
  out.norm = std::sqrt(linalg_innerProd((dlong)NL, xL.data(), MxL.data()));
  out.xL = std::move(xL);
  return out;
}

}  // namespace oracle
