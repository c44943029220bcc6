// Smoothed-aggregation AMG for the p-multigrid base level (see amg.hpp).
#include "amg.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "elliptic.hpp"

namespace libp {

namespace {

// row-wise sum of duplicates after sorting (row, col) triplets
struct triplet_t {
  dlong r, c;
  double v;
};

csrHost_t fromTriplets(std::vector<triplet_t>& t, dlong n, dlong m) {
  std::sort(t.begin(), t.end(), [](const triplet_t& a, const triplet_t& b) {
    return a.r != b.r ? a.r < b.r : a.c < b.c;
  });
  csrHost_t A;
  A.n = n;
  A.m = m;
  A.rs.assign(n + 1, 0);
  for (size_t k = 0; k < t.size();) {
    size_t e = k;
    double v = 0;
    while (e < t.size() && t[e].r == t[k].r && t[e].c == t[k].c) v += t[e++].v;
    A.ci.push_back(t[k].c);
    A.v.push_back(v);
    A.rs[t[k].r + 1]++;
    k = e;
  }
  for (dlong i = 0; i < n; ++i) A.rs[i + 1] += A.rs[i];
  return A;
}

csrHost_t transpose(const csrHost_t& A) {
  csrHost_t T;
  T.n = A.m;
  T.m = A.n;
  T.rs.assign(T.n + 1, 0);
  for (dlong k = 0; k < A.nnz(); ++k) T.rs[A.ci[k] + 1]++;
  for (dlong i = 0; i < T.n; ++i) T.rs[i + 1] += T.rs[i];
  T.ci.resize(A.nnz());
  T.v.resize(A.nnz());
  std::vector<dlong> pos(T.rs.begin(), T.rs.end() - 1);
  for (dlong i = 0; i < A.n; ++i)  // rows in order: the columns of T stay ascending
    for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k) {
      const dlong p = pos[A.ci[k]]++;
      T.ci[p] = i;
      T.v[p] = A.v[k];
    }
  return T;
}

// C = A B (Gustavson, columns sorted per row)
csrHost_t multiply(const csrHost_t& A, const csrHost_t& B) {
  LIBP_ABORT("AMG: matrix shapes do not match", A.m != B.n);
  csrHost_t C;
  C.n = A.n;
  C.m = B.m;
  C.rs.assign(C.n + 1, 0);
  std::vector<double> acc(B.m, 0.0);
  std::vector<dlong> mark(B.m, -1), cols;
  for (dlong i = 0; i < A.n; ++i) {
    cols.clear();
    for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k) {
      const dlong a = A.ci[k];
      const double av = A.v[k];
      for (dlong q = B.rs[a]; q < B.rs[a + 1]; ++q) {
        const dlong c = B.ci[q];
        if (mark[c] != i) {
          mark[c] = i;
          acc[c] = 0.0;
          cols.push_back(c);
        }
        acc[c] += av * B.v[q];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (dlong c : cols) {
      C.ci.push_back(c);
      C.v.push_back(acc[c]);
    }
    C.rs[i + 1] = (dlong)C.ci.size();
  }
  return C;
}

std::vector<double> diagonal(const csrHost_t& A) {
  std::vector<double> d(A.n, 0.0);
  for (dlong i = 0; i < A.n; ++i)
    for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k)
      if (A.ci[k] == i) d[i] = A.v[k];
  return d;
}

// parCSR::rhoDinvA (parAlmondparCSR.cpp:364-460): a 10-step Arnoldi on D^{-1} A
// from an erand48 start (the reference's drand48 stream); rho is the largest
// |eigenvalue| of the Hessenberg matrix
double rhoDinvA(const csrHost_t& A, const std::vector<double>& invD, unsigned short xs[3]) {
  const dlong N = A.n;
  const int k = (int)std::min<dlong>(10, N);
  if (k == 0) return 0.0;
  std::vector<double> H((size_t)k * k, 0.0);
  std::vector<std::vector<double>> V(k + 1, std::vector<double>(N));
  for (auto& v : V[0]) v = erand48(xs);
  double nv = 0.0;
  for (double v : V[0]) nv += v * v;
  nv = std::sqrt(nv);
  for (auto& v : V[0]) v *= 1.0 / nv;
  for (int j = 0; j < k; ++j) {
    std::vector<double>& w = V[j + 1];
    for (dlong i = 0; i < N; ++i) {
      double s = 0.0;
      for (dlong q = A.rs[i]; q < A.rs[i + 1]; ++q) s += A.v[q] * V[j][A.ci[q]];
      w[i] = s * invD[i];
    }
    for (int i = 0; i <= j; ++i) {  // modified Gram-Schmidt
      double hij = 0.0;
      for (dlong n = 0; n < N; ++n) hij += V[i][n] * w[n];
      for (dlong n = 0; n < N; ++n) w[n] -= hij * V[i][n];
      H[i + (size_t)j * k] = hij;
    }
    if (j + 1 < k) {
      double nj = 0.0;
      for (double v : w) nj += v * v;
      nj = std::sqrt(nj);
      H[j + 1 + (size_t)j * k] = nj;
      for (auto& v : w) v *= 1.0 / nj;
    }
  }
  double rho = 0.0;
  for (auto& l : hessenbergEigenvalues(k, H)) rho = std::max(rho, std::abs(l));
  return rho;
}

// (state, weight, id) order of the MIS-2 selection (parAlmondFormAggregates.cpp:35-47)
inline bool greater(int s, double r, dlong i, int smax, double rmax, dlong imax) {
  if (s != smax) return s > smax;
  if (r != rmax) return r > rmax;
  return i > imax;
}

/* Aggregation by a distance-2 maximal independent set on the strong graph C
 * (parAlmondFormAggregates.cpp): weights = column counts + a uniform draw;
 * repeated two-ring sweeps elect the local maxima as roots and remove the
 * nodes within distance 2 of a root; every other node then joins the
 * aggregate of the strongest root in its first, then second, ring. Returns
 * the aggregate of each node and the aggregate count. */
std::vector<dlong> formAggregates(const csrHost_t& C, unsigned short xs[3], dlong& nAgg) {
  const dlong N = C.n;
  std::vector<double> w(N);
  for (dlong i = 0; i < N; ++i) w[i] = erand48(xs);
  std::vector<int> cnt(N, 0);
  for (dlong k = 0; k < C.nnz(); ++k) cnt[C.ci[k]]++;
  for (dlong i = 0; i < N; ++i) w[i] += cnt[i];
  std::vector<int> state(N, 0), Ts(N);
  std::vector<double> Tr(N);
  std::vector<dlong> Ti(N);
  for (;;) {
    for (dlong i = 0; i < N; ++i) {
      int smax = state[i];
      double rmax = w[i];
      dlong imax = i;
      if (smax != 1)
        for (dlong k = C.rs[i]; k < C.rs[i + 1]; ++k) {
          const dlong c = C.ci[k];
          if (c != i && greater(state[c], w[c], c, smax, rmax, imax)) {
            smax = state[c];
            rmax = w[c];
            imax = c;
          }
        }
      Ts[i] = smax;
      Tr[i] = rmax;
      Ti[i] = imax;
    }
    for (dlong i = 0; i < N; ++i) {
      int smax = Ts[i];
      double rmax = Tr[i];
      dlong imax = Ti[i];
      for (dlong k = C.rs[i]; k < C.rs[i + 1]; ++k) {
        const dlong c = C.ci[k];
        if (c != i && greater(Ts[c], Tr[c], Ti[c], smax, rmax, imax)) {
          smax = Ts[c];
          rmax = Tr[c];
          imax = Ti[c];
        }
      }
      if (state[i] == 0 && imax == i) state[i] = 1;
      if (state[i] == 0 && smax == 1) state[i] = -1;
    }
    dlong undecided = 0;
    for (dlong i = 0; i < N; ++i) undecided += (state[i] == 0);
    if (!undecided) break;
  }
  std::vector<dlong> agg(N, -1);
  nAgg = 0;
  for (dlong i = 0; i < N; ++i)
    if (state[i] == 1) agg[i] = nAgg++;
  // first ring, then second ring: join the strongest root's aggregate
  std::vector<dlong> Tc(N);
  for (dlong i = 0; i < N; ++i) {
    int smax = state[i];
    double rmax = w[i];
    dlong imax = i, cmax = agg[i];
    if (smax != 1)
      for (dlong k = C.rs[i]; k < C.rs[i + 1]; ++k) {
        const dlong c = C.ci[k];
        if (c != i && greater(state[c], w[c], c, smax, rmax, imax)) {
          smax = state[c];
          rmax = w[c];
          imax = c;
          cmax = agg[c];
        }
      }
    Ts[i] = smax;
    Tr[i] = rmax;
    Ti[i] = imax;
    Tc[i] = cmax;
  }
  for (dlong i = 0; i < N; ++i)
    if (agg[i] == -1 && Ts[i] == 1 && Tc[i] > -1) agg[i] = Tc[i];
  for (dlong i = 0; i < N; ++i) {
    int smax = Ts[i];
    double rmax = Tr[i];
    dlong imax = Ti[i], cmax = Tc[i];
    for (dlong k = C.rs[i]; k < C.rs[i + 1]; ++k) {
      const dlong c = C.ci[k];
      if (c != i && greater(Ts[c], Tr[c], Ti[c], smax, rmax, imax)) {
        smax = Ts[c];
        rmax = Tr[c];
        imax = Ti[c];
        cmax = Tc[c];
      }
    }
    if (agg[i] == -1 && smax == 1 && cmax > -1) agg[i] = cmax;
  }
  // a node that no root reaches within two rings forms its own aggregate
  for (dlong i = 0; i < N; ++i)
    if (agg[i] == -1) agg[i] = nAgg++;
  return agg;
}

// strength of connection (parAlmondStrongGraph.cpp:39-222), diagonal kept:
// symmetric |a_ij| > theta sqrt(|a_ii a_jj|), or Ruge-Stuben -s_i a_ij > theta max_k(-s_i a_ik)
csrHost_t strongGraph(const csrHost_t& A, const std::vector<double>& d, double theta, bool rugeStuben) {
  csrHost_t C;
  C.n = A.n;
  C.m = A.m;
  C.rs.assign(A.n + 1, 0);
  for (dlong i = 0; i < A.n; ++i) {
    const double aii = std::fabs(d[i]);
    const double sign = (d[i] >= 0) ? 1.0 : -1.0;
    double maxOD = 0.0;
    if (rugeStuben)
      for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k)
        if (A.ci[k] != i) maxOD = std::max(maxOD, -sign * A.v[k]);
    for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k) {
      const dlong j = A.ci[k];
      const bool strong = rugeStuben ? (-sign * A.v[k] > theta * maxOD)
                                     : (std::fabs(A.v[k]) > theta * std::sqrt(aii * std::fabs(d[j])));
      if (j == i || strong) {
        C.ci.push_back(j);
        C.v.push_back(1.0);
      }
    }
    C.rs[i + 1] = (dlong)C.ci.size();
  }
  return C;
}

std::vector<double> inverseDiagonal(const std::vector<double>& d) {
  std::vector<double> inv(d.size());
  for (size_t i = 0; i < d.size(); ++i) inv[i] = (d[i] != 0.0) ? 1.0 / d[i] : 0.0;
  return inv;
}

int lanesFor(const csrHost_t& A) {
  const double avg = A.n ? (double)A.nnz() / A.n : 0.0;
  return avg >= 48 ? 64 : (avg >= 6 ? 8 : 1);
}

}  // namespace

csrHost_t ellipticAssembleCSR(elliptic_t& el) {
  const mesh_t& mesh = el.mesh;
  const dlong E = mesh.Nelements;
  const int Np = mesh.Np;
  const size_t NL = (size_t)E * Np;
  LIBP_ABORT("AMG: assembling degree " << mesh.N << " elements is not supported (N <= 3)", mesh.N > 3);
  platform_t& p = el.platform;
  std::vector<dlong> g2l(NL);
  el.o_GlobalToLocal.copyTo(g2l.data(), NL, 0, p.stream);
  deviceMemory<double> o_u = p.malloc<double>(NL), o_Au = p.malloc<double>(NL);
  std::vector<double> u(NL), Au(NL), Ae((size_t)E * Np * Np);
  for (int n = 0; n < Np; ++n) {
    for (size_t i = 0; i < NL; ++i) u[i] = ((int)(i % Np) == n) ? 1.0 : 0.0;
    o_u.copyFrom(u.data(), NL, 0, p.stream);
    LIBP_CALL(libp_ax_hex_local_f64(mesh.N, E, mesh.o_wJ.ptr(), mesh.o_ggeo.ptr(), mesh.o_D.ptr(), el.lambda,
                                    o_u.ptr(), o_Au.ptr(), p.s()));
    o_Au.copyTo(Au.data(), NL, 0, p.stream);
    p.finish();
    for (dlong e = 0; e < E; ++e)
      for (int a = 0; a < Np; ++a) Ae[((size_t)e * Np + a) * Np + n] = Au[(size_t)e * Np + a];
  }
  std::vector<triplet_t> t;
  t.reserve(NL * Np);
  for (dlong e = 0; e < E; ++e)
    for (int a = 0; a < Np; ++a) {
      const dlong r = g2l[(size_t)e * Np + a];
      if (r < 0) continue;
      for (int b = 0; b < Np; ++b) {
        const dlong c = g2l[(size_t)e * Np + b];
        const double v = Ae[((size_t)e * Np + a) * Np + b];
        // entries below the reference's nonZeroThreshold are dropped (:832-834)
        if (c >= 0 && std::fabs(v) >= 1e-7) t.push_back({r, c, v});
      }
    }
  return fromTriplets(t, el.Ndofs, el.Ndofs);
}

amgSolver_t::options_t amgSolver_t::optionsFrom(const settings_t& settings) {
  options_t o;
  std::string v;
  settings.getSetting("PARALMOND CYCLE", v);
  LIBP_ABORT("PARALMOND CYCLE " << v << " is not supported by this build (VCYCLE only)",
             !v.empty() && v != "VCYCLE");
  o.smoothedAggregation = !settings.compareSetting("PARALMOND AGGREGATION", "UNSMOOTHED");
  o.rugeStuben = settings.compareSetting("PARALMOND STRENGTH", "RUGESTUBEN");
  o.chebyshev = !settings.compareSetting("PARALMOND SMOOTHER", "DAMPEDJACOBI");
  settings.getSetting("PARALMOND CHEBYSHEV DEGREE", o.chebDegree);
  LIBP_ABORT("PARALMOND CHEBYSHEV DEGREE must be at least 1", o.chebyshev && o.chebDegree < 1);
  return o;
}

amgSolver_t::amgSolver_t(elliptic_t& el, unsigned short xs[3]) : platform(el.platform) {
  opt = optionsFrom(el.settings);
  build(ellipticAssembleCSR(el), xs, el.allNeumann != 0, el.allNeumannPenalty);
}

amgSolver_t::amgSolver_t(platform_t& _platform, const csrHost_t& A, unsigned short xs[3], const options_t& o,
                         bool nullSpace, double nullSpacePenalty)
    : platform(_platform), opt(o) {
  build(A, xs, nullSpace, nullSpacePenalty);
}

void amgSolver_t::build(const csrHost_t& A0, unsigned short xs[3], bool nullSpace, double nullSpacePenalty) {
  levels.clear();
  levels.emplace_back();
  levels[0].A = A0;
  // parCSR::diagSetup of the fine matrix (its Arnoldi draws come first)
  levels[0].invDiag = inverseDiagonal(diagonal(A0));
  levels[0].rho = rhoDinvA(A0, levels[0].invDiag, xs);
  null.assign(A0.n, 1.0 / std::sqrt((double)A0.n));  // ellipticPreconMultiGrid.cpp:140-146
  // theta 0.08 halved on every level (symmetric) or 0.5 (Ruge-Stuben), parAlmondAMGSetup.cpp:86-119
  double theta = opt.rugeStuben ? 0.5 : 0.08;
  dlong n = A0.n;
  while (n > targetSize) {
    level_t& L = levels.back();
    const csrHost_t& A = L.A;
    // amgLevel::setupSmoother (parAlmondAMGLevel.cpp:76-84)
    L.lambda1 = L.rho;
    L.lambda0 = L.rho / 10.;
    L.lambda = (4. / 3.) / L.rho;
    // strength, aggregates, tentative prolongator with normalised columns
    const std::vector<double> d = diagonal(A);
    const csrHost_t C = strongGraph(A, d, theta, opt.rugeStuben);
    dlong nc = 0;
    L.aggregate = formAggregates(C, xs, nc);
    const std::vector<dlong>& agg = L.aggregate;
    std::vector<double> nullc(nc, 0.0);
    for (dlong i = 0; i < A.n; ++i) nullc[agg[i]] += null[i] * null[i];
    for (auto& v : nullc) v = std::sqrt(v);
    csrHost_t T;
    T.n = A.n;
    T.m = nc;
    T.rs.resize(A.n + 1);
    T.ci.resize(A.n);
    T.v.resize(A.n);
    for (dlong i = 0; i < A.n; ++i) {
      T.rs[i] = i;
      T.ci[i] = agg[i];
      T.v[i] = null[i] / nullc[agg[i]];
    }
    T.rs[A.n] = A.n;
    csrHost_t Ac;
    if (opt.smoothedAggregation) {
      // P = T - omega D^{-1} A T, omega = 4/3 / rho; Ac = R (A P)
      csrHost_t AT = multiply(A, T);
      const double omega = (4. / 3.) / L.rho;
      std::vector<triplet_t> t;
      for (dlong i = 0; i < A.n; ++i) {
        t.push_back({i, T.ci[i], T.v[i]});
        for (dlong k = AT.rs[i]; k < AT.rs[i + 1]; ++k) t.push_back({i, AT.ci[k], -omega * L.invDiag[i] * AT.v[k]});
      }
      L.P = fromTriplets(t, A.n, nc);
      L.R = transpose(L.P);
      Ac = multiply(L.R, multiply(A, L.P));
    } else {
      // P = T; (T^T A T)_IJ = sum over i in I, j in J of T_iI a_ij T_jJ (parAlmondGalerkinProd.cpp)
      L.P = T;
      L.R = transpose(T);
      std::vector<triplet_t> t;
      t.reserve(A.nnz());
      for (dlong i = 0; i < A.n; ++i)
        for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k)
          t.push_back({agg[i], agg[A.ci[k]], A.v[k] * T.v[i] * T.v[A.ci[k]]});
      Ac = fromTriplets(t, nc, nc);
    }
    levels.emplace_back();
    level_t& Lc = levels.back();
    Lc.A = std::move(Ac);
    Lc.invDiag = inverseDiagonal(diagonal(Lc.A));
    Lc.rho = rhoDinvA(Lc.A, Lc.invDiag, xs);
    null = std::move(nullc);
    if (!opt.rugeStuben) theta *= 0.5;
    const dlong ncNew = Lc.A.n;
    if (ncNew <= targetSize || n < 2 * ncNew) break;  // coarse enough, or coarsening stalls
    n = ncNew;
  }
  // coarsest level: dense inverse by Gauss-Jordan with partial pivoting (parAlmondCoarseExact.cpp)
  const csrHost_t& Ac = levels.back().A;
  const dlong nc = Ac.n;
  std::vector<double> M((size_t)nc * nc, 0.0), I((size_t)nc * nc, 0.0);
  for (dlong i = 0; i < nc; ++i) {
    I[(size_t)i * nc + i] = 1.0;
    for (dlong k = Ac.rs[i]; k < Ac.rs[i + 1]; ++k) M[(size_t)i * nc + Ac.ci[k]] = Ac.v[k];
  }
  if (nullSpace)  // exactSolver_t::setup, parAlmondCoarseExact.cpp:154-160
    for (dlong i = 0; i < nc; ++i)
      for (dlong j = 0; j < nc; ++j) M[(size_t)i * nc + j] += nullSpacePenalty * null[i] * null[j];
  for (dlong c = 0; c < nc; ++c) {
    dlong piv = c;
    for (dlong r = c + 1; r < nc; ++r)
      if (std::fabs(M[(size_t)r * nc + c]) > std::fabs(M[(size_t)piv * nc + c])) piv = r;
    LIBP_ABORT("AMG: singular coarsest matrix", M[(size_t)piv * nc + c] == 0.0);
    if (piv != c)
      for (dlong j = 0; j < nc; ++j) {
        std::swap(M[(size_t)c * nc + j], M[(size_t)piv * nc + j]);
        std::swap(I[(size_t)c * nc + j], I[(size_t)piv * nc + j]);
      }
    const double inv = 1.0 / M[(size_t)c * nc + c];
    for (dlong j = 0; j < nc; ++j) {
      M[(size_t)c * nc + j] *= inv;
      I[(size_t)c * nc + j] *= inv;
    }
    for (dlong r = 0; r < nc; ++r) {
      if (r == c) continue;
      const double f = M[(size_t)r * nc + c];
      if (f == 0.0) continue;
      for (dlong j = 0; j < nc; ++j) {
        M[(size_t)r * nc + j] -= f * M[(size_t)c * nc + j];
        I[(size_t)r * nc + j] -= f * I[(size_t)c * nc + j];
      }
    }
  }
  o_coarseInv = platform.malloc(I);
  // device arrays
  for (size_t l = 0; l < levels.size(); ++l) {
    level_t& L = levels[l];
    const dlong n = L.A.n;
    L.o_Ars = platform.malloc(L.A.rs);
    L.o_Aci = platform.malloc(L.A.ci);
    L.o_Av = platform.malloc(L.A.v);
    L.o_invDiag = platform.malloc(L.invDiag);
    L.lanesA = lanesFor(L.A);
    if (l + 1 < levels.size()) {
      L.o_Prs = platform.malloc(L.P.rs);
      L.o_Pci = platform.malloc(L.P.ci);
      L.o_Pv = platform.malloc(L.P.v);
      L.o_Rrs = platform.malloc(L.R.rs);
      L.o_Rci = platform.malloc(L.R.ci);
      L.o_Rv = platform.malloc(L.R.v);
      L.lanesP = lanesFor(L.P);
      L.lanesR = lanesFor(L.R);
    }
    L.o_b = platform.malloc<double>(std::max<dlong>(n, 1));
    L.o_x = platform.malloc<double>(std::max<dlong>(n, 1));
    L.o_res = platform.malloc<double>(std::max<dlong>(n, 1));
    L.o_d = platform.malloc<double>(std::max<dlong>(n, 1));
    L.o_Ad = platform.malloc<double>(std::max<dlong>(n, 1));
  }
  platform.finish();
}

void amgSolver_t::spmv(const deviceMemory<dlong>& rs, const deviceMemory<dlong>& ci, const deviceMemory<double>& v,
                       dlong n, int lanes, double alpha, const double* x, double beta, const double* z, double* y) {
  LIBP_CALL(libp_csr_spmv_f64(n, rs.ptr(), ci.ptr(), v.ptr(), alpha, x, beta, z, y, lanes, platform.s()));
}

// amgLevel::smooth (parAlmondAMGLevel.cpp:63-72): parCSR::smoothChebyshev
// (parAlmondAMGSmoother.cpp:68-160) with the vector updates fused as in the p-MG
// smoother, or parCSR::smoothDampedJacobi (:36-66)
void amgSolver_t::smooth(size_t l, const double* b, double* x, bool xIsZero) {
  level_t& L = levels[l];
  const dlong N = L.A.n;
  void* s = platform.s();
  if (!opt.chebyshev) {
    if (xIsZero) {
      LIBP_CALL(libp_amxpy_f64(N, L.lambda, L.o_invDiag.ptr(), b, 0.0, x, s));
      return;
    }
    // x += lambda D^{-1} (b - A x)
    spmv(L.o_Ars, L.o_Aci, L.o_Av, N, L.lanesA, -1.0, x, 1.0, b, L.o_res.ptr());
    LIBP_CALL(libp_amxpy_f64(N, L.lambda, L.o_invDiag.ptr(), L.o_res.ptr(), 1.0, x, s));
    return;
  }
  const double theta = 0.5 * (L.lambda1 + L.lambda0), delta = 0.5 * (L.lambda1 - L.lambda0);
  const double sigma = theta / delta;
  double rho_n = 1. / sigma;
  if (xIsZero) {
    LIBP_CALL(libp_cheb_start_f64(N, L.o_invDiag.ptr(), b, nullptr, 1.0 / theta, L.o_res.ptr(), L.o_d.ptr(), x, 1, s));
  } else {
    spmv(L.o_Ars, L.o_Aci, L.o_Av, N, L.lanesA, 1.0, x, 0.0, nullptr, L.o_Ad.ptr());
    LIBP_CALL(libp_cheb_start_f64(N, L.o_invDiag.ptr(), b, L.o_Ad.ptr(), 1.0 / theta, L.o_res.ptr(), L.o_d.ptr(), x,
                                  0, s));
  }
  for (int k = 0; k < opt.chebDegree; ++k) {
    spmv(L.o_Ars, L.o_Aci, L.o_Av, N, L.lanesA, 1.0, L.o_d.ptr(), 0.0, nullptr, L.o_Ad.ptr());
    const double rho_np1 = 1.0 / (2. * sigma - rho_n);
    LIBP_CALL(libp_cheb_step_f64(N, L.o_invDiag.ptr(), L.o_Ad.ptr(), 2.0 * rho_np1 / delta, rho_np1 * rho_n,
                                 L.o_res.ptr(), L.o_d.ptr(), x, s));
    rho_n = rho_np1;
  }
}

// parAlmondVcycle.cpp:34-60
void amgSolver_t::cycle(size_t l, const double* b, double* x) {
  level_t& L = levels[l];
  if (l + 1 == levels.size()) {
    LIBP_CALL(libp_dense_gemv_f64(L.A.n, o_coarseInv.ptr(), b, x, platform.s()));
    return;
  }
  level_t& Lc = levels[l + 1];
  smooth(l, b, x, true);
  // res = b - A x ; bc = R res
  spmv(L.o_Ars, L.o_Aci, L.o_Av, L.A.n, L.lanesA, -1.0, x, 1.0, b, L.o_res.ptr());
  spmv(L.o_Rrs, L.o_Rci, L.o_Rv, L.R.n, L.lanesR, 1.0, L.o_res.ptr(), 0.0, nullptr, Lc.o_b.ptr());
  cycle(l + 1, Lc.o_b.ptr(), Lc.o_x.ptr());
  // x += P xc
  spmv(L.o_Prs, L.o_Pci, L.o_Pv, L.P.n, L.lanesP, 1.0, Lc.o_x.ptr(), 1.0, x, x);
  smooth(l, b, x, false);
}

void amgSolver_t::vcycle(const double* b, double* x) { cycle(0, b, x); }

}  // namespace libp
