// CPU oracle (test infrastructure only): the parAlmond algebraic multigrid that
// ends the reference's p-multigrid (libs/parAlmond), restated for one rank, and
// the assembled continuous operator it is built on. Every MPI exchange of the
// reference is a size-1 no-op here and is dropped.
#include <algorithm>
#include <cmath>
#include <complex>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

std::vector<std::complex<double>> hessenbergEigs(int n, const std::vector<double>& Hcm);

namespace {
struct Coo {
  hlong row, col;
  double val;
};
bool cooLess(const Coo& a, const Coo& b) { return a.row != b.row ? a.row < b.row : a.col < b.col; }

// parCSR(parCOO) on one rank (parAlmondparCSR.cpp): entries sorted by (row, col),
// duplicates summed in that order
CsrO fromCoo(std::vector<Coo>& e, dlong n, dlong m) {
  std::sort(e.begin(), e.end(), cooLess);
  CsrO A;
  A.n = n;
  A.m = m;
  A.rs.assign(n + 1, 0);
  size_t k = 0;
  while (k < e.size()) {
    size_t j = k;
    double v = e[k].val;
    while (j + 1 < e.size() && e[j + 1].row == e[k].row && e[j + 1].col == e[k].col) v += e[++j].val;
    A.ci.push_back((dlong)e[k].col);
    A.v.push_back(v);
    A.rs[e[k].row + 1]++;
    k = j + 1;
  }
  for (dlong i = 0; i < n; ++i) A.rs[i + 1] += A.rs[i];
  return A;
}

// C = A B, rows formed in A's column order, duplicates summed after sorting
// (parAlmondSpMM.cpp builds the product by rows and sorts each row's entries)
CsrO spmm(const CsrO& A, const CsrO& B) {
  std::vector<Coo> e;
  for (dlong i = 0; i < A.n; ++i)
    for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k) {
      const dlong a = A.ci[k];
      for (dlong q = B.rs[a]; q < B.rs[a + 1]; ++q) e.push_back({i, B.ci[q], A.v[k] * B.v[q]});
    }
  return fromCoo(e, A.n, B.m);
}

// parAlmondTranspose.cpp
CsrO transposeCsr(const CsrO& A) {
  std::vector<Coo> e;
  for (dlong i = 0; i < A.n; ++i)
    for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k) e.push_back({A.ci[k], i, A.v[k]});
  return fromCoo(e, A.m, A.n);
}
}  // namespace

void CsrO::spmv(double alpha, const double* x, double beta, double* y) const {
  for (dlong i = 0; i < n; ++i) {
    double s = 0.0;
    for (dlong k = rs[i]; k < rs[i + 1]; ++k) s += v[k] * x[ci[k]];
    y[i] = (beta != 0.0) ? alpha * s + beta * y[i] : alpha * s;
  }
}

/* elliptic_t::BuildOperatorMatrixContinuousHex3D
 * (solvers/elliptic/src/ellipticBuildOperatorMatrixContinuous.cpp:697-880) on one
 * rank: the element stiffness from ggeo and D plus lambda*wJ on the diagonal,
 * between unmasked nodes (maskedGlobalNumbering = GlobalToLocal on one rank),
 * entries with |val| >= 1e-7, sorted by (row, col) with duplicates summed. */
CsrO BuildOperatorMatrixContinuousHex3D(const Elliptic& el) {
  const Mesh& mesh = el.mesh;
  const int Nq = mesh.Nq, Np = mesh.Np;
  const std::vector<double>& D = mesh.D;
  std::vector<Coo> e;
  for (dlong el_ = 0; el_ < mesh.Nelements; ++el_) {
    const double* G = mesh.ggeo.data() + (size_t)el_ * Np * 6;
    auto g = [&](int gid, int id) { return G[id + gid * Np]; };
    for (int nz = 0; nz < Nq; nz++)
      for (int ny = 0; ny < Nq; ny++)
        for (int nx = 0; nx < Nq; nx++) {
          const int idn = nx + ny * Nq + nz * Nq * Nq;
          const dlong row = el.GlobalToLocal[(size_t)el_ * Np + idn];
          if (row < 0) continue;
          for (int mz = 0; mz < Nq; mz++)
            for (int my = 0; my < Nq; my++)
              for (int mx = 0; mx < Nq; mx++) {
                const int idm = mx + my * Nq + mz * Nq * Nq;
                const dlong col = el.GlobalToLocal[(size_t)el_ * Np + idm];
                if (col < 0) continue;
                double val = 0.;
                if ((ny == my) && (nz == mz))
                  for (int k = 0; k < Nq; k++) val += g(0, k + ny * Nq + nz * Nq * Nq) * D[nx + k * Nq] * D[mx + k * Nq];
                if (nz == mz) {
                  val += g(1, mx + ny * Nq + nz * Nq * Nq) * D[nx + mx * Nq] * D[my + ny * Nq];
                  val += g(1, nx + my * Nq + nz * Nq * Nq) * D[mx + nx * Nq] * D[ny + my * Nq];
                }
                if (ny == my) {
                  val += g(2, mx + ny * Nq + nz * Nq * Nq) * D[nx + mx * Nq] * D[mz + nz * Nq];
                  val += g(2, nx + ny * Nq + mz * Nq * Nq) * D[mx + nx * Nq] * D[nz + mz * Nq];
                }
                if ((nx == mx) && (nz == mz))
                  for (int k = 0; k < Nq; k++) val += g(3, nx + k * Nq + nz * Nq * Nq) * D[ny + k * Nq] * D[my + k * Nq];
                if (nx == mx) {
                  val += g(4, nx + my * Nq + nz * Nq * Nq) * D[ny + my * Nq] * D[mz + nz * Nq];
                  val += g(4, nx + ny * Nq + mz * Nq * Nq) * D[my + ny * Nq] * D[nz + mz * Nq];
                }
                if ((nx == mx) && (ny == my))
                  for (int k = 0; k < Nq; k++) val += g(5, nx + ny * Nq + k * Nq * Nq) * D[nz + k * Nq] * D[mz + k * Nq];
                if ((nx == mx) && (ny == my) && (nz == mz)) val += mesh.wJ[(size_t)el_ * Np + idn] * el.lambda;
                if (std::fabs(val) >= 1e-7) e.push_back({row, col, val});
              }
        }
  }
  return fromCoo(e, el.Ndofs, el.Ndofs);
}

/* parCSR::rhoDinvA (parAlmondparCSR.cpp:364-460): 10-step Arnoldi on D^-1 A from
 * a drand48 start, the largest |eigenvalue| of the Hessenberg matrix. */
static double rhoDinvA(const CsrO& A, const std::vector<double>& diagInv, Drand48& drng) {
  const dlong N = A.n;
  int k = 10;
  if (k > N) k = N;
  std::vector<double> H((size_t)k * k, 0.0), Vx(N);
  std::vector<std::vector<double>> V(k + 1, std::vector<double>(N));
  for (dlong n = 0; n < N; n++) Vx[n] = drng.next();
  double norm_vo = 0.0;
  for (dlong n = 0; n < N; n++) norm_vo += Vx[n] * Vx[n];
  norm_vo = std::sqrt(norm_vo);
  for (dlong n = 0; n < N; n++) Vx[n] *= (1.0 / norm_vo);
  V[0] = Vx;
  for (int j = 0; j < k; j++) {
    Vx = V[j];
    A.spmv(1.0, Vx.data(), 0.0, V[j + 1].data());
    for (dlong n = 0; n < N; n++) V[j + 1][n] *= diagInv[n];
    for (int i = 0; i <= j; i++) {
      double hij = 0.0;
      for (dlong n = 0; n < N; n++) hij += V[i][n] * V[j + 1][n];
      for (dlong n = 0; n < N; n++) V[j + 1][n] += -hij * V[i][n];
      H[i + j * k] = hij;
    }
    if (j + 1 < k) {
      double norm_vj = 0.0;
      for (dlong n = 0; n < N; n++) norm_vj += V[j + 1][n] * V[j + 1][n];
      norm_vj = std::sqrt(norm_vj);
      H[j + 1 + j * k] = norm_vj;
      for (dlong n = 0; n < N; n++) V[j + 1][n] *= (1. / H[j + 1 + j * k]);
    }
  }
  double RHO = 0.;
  for (auto& l : hessenbergEigs(k, H)) RHO = std::max(RHO, std::abs(l));
  return RHO;
}

// parCSR::diagSetup (parAlmondparCSR.cpp:330-356)
static void diagSetup(AmgLevelO& L, Drand48& drng) {
  const CsrO& A = L.A;
  L.diagA.assign(A.n, 0.0);
  for (dlong i = 0; i < A.n; i++)
    for (dlong j = A.rs[i]; j < A.rs[i + 1]; j++)
      if (A.ci[j] == i) L.diagA[i] = A.v[j];
  L.diagInv.resize(A.n);
  for (dlong n = 0; n < A.n; n++) L.diagInv[n] = (L.diagA[n] != 0.0) ? 1.0 / L.diagA[n] : 0.0;
  L.rho = rhoDinvA(A, L.diagInv, drng);
}

/* strongGraph (parAlmondStrongGraph.cpp:39-222): the diagonal plus the strong
 * off-diagonal entries of each row, in the row's column order. */
static void strongGraph(const AmgLevelO& L, bool rugeStuben, double theta, std::vector<dlong>& rs,
                        std::vector<dlong>& ci) {
  const CsrO& A = L.A;
  rs.assign(A.n + 1, 0);
  ci.clear();
  for (dlong i = 0; i < A.n; i++) {
    if (rugeStuben) {  // RugeStubenStrength :48-133
      const int sign = (L.diagA[i] >= 0) ? 1 : -1;
      double maxOD = 0.0;
      for (dlong jj = A.rs[i]; jj < A.rs[i + 1]; jj++) {
        if (A.ci[jj] == i) continue;
        const double OD = -sign * A.v[jj];
        if (OD > maxOD) maxOD = OD;
      }
      for (dlong jj = A.rs[i]; jj < A.rs[i + 1]; jj++) {
        const dlong col = A.ci[jj];
        if (col == i) {
          ci.push_back(col);
          continue;
        }
        if (-sign * A.v[jj] > theta * maxOD) ci.push_back(col);
      }
    } else {  // SymmetricStrength :135-222
      const double Aii = std::fabs(L.diagA[i]);
      for (dlong jj = A.rs[i]; jj < A.rs[i + 1]; jj++) {
        const dlong col = A.ci[jj];
        if (col == i) {
          ci.push_back(col);
          continue;
        }
        const double Ajj = std::fabs(L.diagA[col]);
        if (std::fabs(A.v[jj]) > theta * (std::sqrt(Aii * Ajj))) ci.push_back(col);
      }
    }
    rs[i + 1] = (dlong)ci.size();
  }
}

// customLess (parAlmondFormAggregates.cpp:35-48)
static bool customLess(int smax, double rmax, hlong imax, int s, double r, hlong i) {
  if (s > smax) return true;
  if (smax > s) return false;
  if (r > rmax) return true;
  if (rmax > r) return false;
  if (i > imax) return true;
  if (i < imax) return false;
  return false;
}

/* formAggregates (parAlmondFormAggregates.cpp:56-288) on one rank: distance-2
 * maximal independent set by (state, drand48 + column count, id), then every node
 * joins the strongest root within its first, then second, ring. The loops run in
 * row order, as the reference's (its OpenMP pragmas are commented out). */
static dlong formAggregates(const std::vector<dlong>& Crs, const std::vector<dlong>& Cci, dlong N, Drand48& drng,
                            std::vector<hlong>& FineToCoarse, std::vector<double>& rands, std::vector<int>& states) {
  rands.assign(N, 0.0);
  states.assign(N, 0);
  std::vector<double> Tr(N);
  std::vector<int> Ts(N);
  std::vector<hlong> Ti(N), Tc(N);
  for (dlong i = 0; i < N; i++) rands[i] = drng.next();
  std::vector<int> colCnt(N, 0);
  for (size_t i = 0; i < Cci.size(); i++) colCnt[Cci[i]]++;
  for (dlong i = 0; i < N; ++i) rands[i] += colCnt[i];
  bool done = false;
  while (!done) {
    for (dlong i = 0; i < N; i++) {
      int smax = states[i];
      double rmax = rands[i];
      hlong imax = i;
      if (smax != 1)
        for (dlong jj = Crs[i]; jj < Crs[i + 1]; jj++) {
          const dlong col = Cci[jj];
          if (col == i) continue;
          if (customLess(smax, rmax, imax, states[col], rands[col], col)) {
            smax = states[col];
            rmax = rands[col];
            imax = col;
          }
        }
      Ts[i] = smax;
      Tr[i] = rmax;
      Ti[i] = imax;
    }
    for (dlong i = 0; i < N; i++) {
      int smax = Ts[i];
      double rmax = Tr[i];
      hlong imax = Ti[i];
      for (dlong jj = Crs[i]; jj < Crs[i + 1]; jj++) {
        const dlong col = Cci[jj];
        if (col == i) continue;
        if (customLess(smax, rmax, imax, Ts[col], Tr[col], Ti[col])) {
          smax = Ts[col];
          rmax = Tr[col];
          imax = Ti[col];
        }
      }
      if ((states[i] == 0) && (imax == i)) states[i] = 1;
      if ((states[i] == 0) && (smax == 1)) states[i] = -1;
    }
    dlong undecided = 0;
    for (dlong n = 0; n < N; n++)
      if (states[n] == 0) undecided++;
    done = (undecided == 0);
  }
  dlong numAggs = 0;
  FineToCoarse.assign(N, -1);
  for (dlong i = 0; i < N; i++)
    if (states[i] == 1) FineToCoarse[i] = numAggs++;
  for (dlong i = 0; i < N; i++) {
    int smax = states[i];
    double rmax = rands[i];
    hlong imax = i;
    hlong cmax = FineToCoarse[i];
    if (smax != 1)
      for (dlong jj = Crs[i]; jj < Crs[i + 1]; jj++) {
        const dlong col = Cci[jj];
        if (col == i) continue;
        if (customLess(smax, rmax, imax, states[col], rands[col], col)) {
          smax = states[col];
          rmax = rands[col];
          imax = col;
          cmax = FineToCoarse[col];
        }
      }
    Ts[i] = smax;
    Tr[i] = rmax;
    Ti[i] = imax;
    Tc[i] = cmax;
    if ((FineToCoarse[i] == -1) && (smax == 1) && (cmax > -1)) FineToCoarse[i] = cmax;
  }
  for (dlong i = 0; i < N; i++) {
    int smax = Ts[i];
    double rmax = Tr[i];
    hlong imax = Ti[i];
    hlong cmax = Tc[i];
    for (dlong jj = Crs[i]; jj < Crs[i + 1]; jj++) {
      const dlong col = Cci[jj];
      if (col == i) continue;
      if (customLess(smax, rmax, imax, Ts[col], Tr[col], Ti[col])) {
        smax = Ts[col];
        rmax = Tr[col];
        imax = Ti[col];
        cmax = Tc[col];
      }
    }
    if ((FineToCoarse[i] == -1) && (smax == 1) && (cmax > -1)) FineToCoarse[i] = cmax;
  }
  return numAggs;
}

/* AMG setup (parAlmondAMGSetup.cpp:35-144 with coarsenAmgLevel,
 * parAlmondCoarsenLevel.cpp:34-88): levels until the coarse size is at most
 * the exact solver's 1000 rows or coarsening stalls (n < 2 nc); theta 0.08
 * halved per level (SYMMETRIC) or 0.5 fixed (RUGESTUBEN); the exact coarse
 * solve (parAlmondCoarseExact.cpp) is a dense inverse, with the null-space
 * penalty added when the operator is singular. */
void AmgO::Setup(const CsrO& A0, const Settings& s, Drand48& drng, bool nullSpace, double nullSpacePenalty) {
  if (s.compare("PARALMOND CYCLE", "KCYCLE") || s.compare("PARALMOND CYCLE", "EXACT"))
    throw std::runtime_error("oracle AMG: only PARALMOND CYCLE VCYCLE is restated");
  const bool rugeStuben = !s.compare("PARALMOND STRENGTH", "SYMMETRIC");
  const bool unsmoothed = s.compare("PARALMOND AGGREGATION", "UNSMOOTHED");
  const bool cheb = s.compare("PARALMOND SMOOTHER", "CHEBYSHEV");
  const int chebIts = s.has("PARALMOND CHEBYSHEV DEGREE") ? s.geti("PARALMOND CHEBYSHEV DEGREE") : 2;
  levels.clear();
  levels.emplace_back();
  levels[0].A = A0;
  diagSetup(levels[0], drng);
  const hlong total = A0.n;
  null.assign(A0.n, 1.0 / std::sqrt((double)total));  // ellipticPreconMultiGrid.cpp:140-146
  const int targetSize = 1000;
  hlong globalSize = A0.n;
  double theta = rugeStuben ? 0.5 : 0.08;
  if (globalSize > targetSize) {
    for (;;) {
      AmgLevelO& L = levels.back();
      // amgLevel::setupSmoother (parAlmondAMGLevel.cpp:76-84)
      L.stype = cheb ? 1 : 0;
      L.ChebyshevIterations = chebIts;
      if (cheb) {
        L.lambda1 = L.rho;
        L.lambda0 = L.rho / 10.;
      } else {
        L.lambda = (4. / 3.) / L.rho;
      }
      strongGraph(L, rugeStuben, theta, L.Crs, L.Cci);
      const dlong nAgg = formAggregates(L.Crs, L.Cci, L.A.n, drng, L.FineToCoarse, L.weights, L.states);
      // tentativeProlongator (parAlmondTentativeProlongator.cpp:34-92)
      std::vector<Coo> e;
      for (dlong n = 0; n < L.A.n; n++) e.push_back({n, L.FineToCoarse[n], null[n]});
      CsrO T = fromCoo(e, L.A.n, nAgg);
      std::vector<double> nullc(nAgg, 0.0);
      for (dlong i = 0; i < T.nnz(); i++) nullc[T.ci[i]] += T.v[i] * T.v[i];
      for (dlong i = 0; i < nAgg; i++) nullc[i] = std::sqrt(nullc[i]);
      for (dlong i = 0; i < T.nnz(); i++) T.v[i] /= nullc[T.ci[i]];
      CsrO P;
      if (!unsmoothed) {
        // smoothProlongator (parAlmondSmoothPrologator.cpp): P = (I - omega D^-1 A) T, omega = 4/3 / rho
        const double omega = (4. / 3.) / L.rho;
        CsrO AT = spmm(L.A, T);
        std::vector<Coo> pe;
        for (dlong i = 0; i < T.n; i++) {
          for (dlong k = T.rs[i]; k < T.rs[i + 1]; k++) pe.push_back({i, T.ci[k], T.v[k]});
          for (dlong k = AT.rs[i]; k < AT.rs[i + 1]; k++) pe.push_back({i, AT.ci[k], -omega * L.diagInv[i] * AT.v[k]});
        }
        P = fromCoo(pe, T.n, T.m);
      } else {
        P = T;
      }
      L.P = P;
      L.R = transposeCsr(P);
      AmgLevelO C;
      if (!unsmoothed) {
        C.A = spmm(L.R, spmm(L.A, P));
      } else {
        // galerkinProd (parAlmondGalerkinProd.cpp): (P^T A P)_IJ = sum P_iI A_ij P_jJ
        std::vector<Coo> ge;
        for (dlong i = 0; i < L.A.n; i++)
          for (dlong j = L.A.rs[i]; j < L.A.rs[i + 1]; j++) {
            const dlong col = L.A.ci[j];
            ge.push_back({L.FineToCoarse[i], L.FineToCoarse[col], L.A.v[j] * P.v[i] * P.v[col]});
          }
        C.A = fromCoo(ge, nAgg, nAgg);
      }
      diagSetup(C, drng);
      levels.push_back(std::move(C));
      null = nullc;
      if (!rugeStuben) theta = theta / 2;
      const hlong globalCoarseSize = levels.back().A.n;
      if (globalCoarseSize <= targetSize || globalSize < 2 * globalCoarseSize) break;
      globalSize = globalCoarseSize;
    }
  }
  // exactSolver_t::setup (parAlmondCoarseExact.cpp:88-190)
  const CsrO& Ac = levels.back().A;
  const int n = Ac.n;
  std::vector<double> M((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (dlong k = Ac.rs[i]; k < Ac.rs[i + 1]; ++k) M[(size_t)i * n + Ac.ci[k]] = Ac.v[k];
  if (nullSpace)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) M[(size_t)i * n + j] += nullSpacePenalty * null[i] * null[j];
  coarseInv = denseInverse(n, M);
}

// Gauss-Jordan elimination with partial pivoting (linAlg_t::matrixInverse in
// the reference calls LAPACK dgetrf/dgetri)
std::vector<double> denseInverse(int n, std::vector<double> M) {
  std::vector<double> I((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(M[(size_t)r * n + c]) > std::fabs(M[(size_t)piv * n + c])) piv = r;
    if (M[(size_t)piv * n + c] == 0.0) throw std::runtime_error("denseInverse: singular matrix");
    if (piv != c)
      for (int j = 0; j < n; ++j) {
        std::swap(M[(size_t)c * n + j], M[(size_t)piv * n + j]);
        std::swap(I[(size_t)c * n + j], I[(size_t)piv * n + j]);
      }
    const double inv = 1.0 / M[(size_t)c * n + c];
    for (int j = 0; j < n; ++j) {
      M[(size_t)c * n + j] *= inv;
      I[(size_t)c * n + j] *= inv;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = M[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        M[(size_t)r * n + j] -= f * M[(size_t)c * n + j];
        I[(size_t)r * n + j] -= f * I[(size_t)c * n + j];
      }
    }
  }
  return I;
}

/* amgLevel::smooth (parAlmondAMGLevel.cpp:63-72): damped Jacobi
 * (parAlmondAMGSmoother.cpp:36-66) or Chebyshev (:68-160). */
void AmgO::smooth(int k, const double* b, double* x, bool xIsZero) const {
  const AmgLevelO& L = levels[k];
  const dlong N = L.A.n;
  std::vector<double> Ax(N);
  if (L.stype == 0) {
    if (xIsZero) {
      for (dlong n = 0; n < N; ++n) x[n] = L.lambda * L.diagInv[n] * b[n];
      return;
    }
    L.A.spmv(1.0, x, 0.0, Ax.data());
    for (dlong n = 0; n < N; ++n) x[n] += L.lambda * L.diagInv[n] * (b[n] - Ax[n]);
    return;
  }
  const double theta = 0.5 * (L.lambda1 + L.lambda0), delta = 0.5 * (L.lambda1 - L.lambda0);
  const double invTheta = 1.0 / theta, sigma = theta / delta;
  double rho_n = 1. / sigma, rho_np1;
  std::vector<double> r(N), d(N);
  if (xIsZero) {
    for (dlong n = 0; n < N; ++n) {
      r[n] = L.diagInv[n] * b[n];
      d[n] = invTheta * r[n];
      x[n] = d[n];
    }
  } else {
    L.A.spmv(1.0, x, 0.0, Ax.data());
    for (dlong n = 0; n < N; ++n) {
      r[n] = L.diagInv[n] * (b[n] - Ax[n]);
      d[n] = invTheta * r[n];
      x[n] += d[n];
    }
  }
  for (int k2 = 0; k2 < L.ChebyshevIterations; k2++) {
    L.A.spmv(1.0, d.data(), 0.0, Ax.data());
    for (dlong n = 0; n < N; ++n) r[n] -= L.diagInv[n] * Ax[n];
    rho_np1 = 1.0 / (2. * sigma - rho_n);
    for (dlong n = 0; n < N; ++n) {
      d[n] = rho_np1 * rho_n * d[n] + 2.0 * rho_np1 / delta * r[n];
      x[n] += d[n];
    }
    rho_n = rho_np1;
  }
}

/* multigrid_t::vcycle (parAlmondVcycle.cpp:34-60) over the AMG levels, the
 * base level by the dense inverse (exactSolver_t::solve, dGEMV). */
void AmgO::vcycle(int k, const double* b, double* x) const {
  const int base = (int)levels.size() - 1;
  if (k == base) {
    const int n = levels[base].A.n;
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += coarseInv[(size_t)i * n + j] * b[j];
      x[i] = s;
    }
    return;
  }
  const AmgLevelO& L = levels[k];
  const dlong N = L.A.n, Nc = levels[k + 1].A.n;
  std::vector<double> res(N), bc(Nc), xc(Nc);
  smooth(k, b, x, true);
  L.A.spmv(1.0, x, 0.0, res.data());
  for (dlong n = 0; n < N; ++n) res[n] = b[n] - res[n];
  L.R.spmv(1.0, res.data(), 0.0, bc.data());
  vcycle(k + 1, bc.data(), xc.data());
  L.P.spmv(1.0, xc.data(), 1.0, x);
  smooth(k, b, x, false);
}

}  // namespace oracle
