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

// spectral radius of D^{-1} A by power iteration (A SPD: D^{-1} A is similar to
// D^{-1/2} A D^{-1/2}); the start vector comes from the erand48 stream
double rhoDinvA(const csrHost_t& A, const std::vector<double>& invD, unsigned short xs[3]) {
  std::vector<double> x(A.n), y(A.n);
  for (auto& v : x) v = erand48(xs);
  double rho = 0;
  for (int it = 0; it < 40; ++it) {
    double nx = 0;
    for (double v : x) nx += v * v;
    nx = std::sqrt(nx);
    if (!(nx > 0)) break;
    for (dlong i = 0; i < A.n; ++i) {
      double s = 0;
      for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k) s += A.v[k] * x[A.ci[k]];
      y[i] = invD[i] * s / nx;
    }
    double ny = 0;
    for (double v : y) ny += v * v;
    rho = std::sqrt(ny);
    x.swap(y);
  }
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

// symmetric strength (parAlmondStrongGraph.cpp:147-230): |a_ij| > theta sqrt(|a_ii a_jj|), diagonal kept
csrHost_t strongGraph(const csrHost_t& A, const std::vector<double>& d, double theta) {
  csrHost_t C;
  C.n = A.n;
  C.m = A.m;
  C.rs.assign(A.n + 1, 0);
  for (dlong i = 0; i < A.n; ++i) {
    const double aii = std::fabs(d[i]);
    for (dlong k = A.rs[i]; k < A.rs[i + 1]; ++k) {
      const dlong j = A.ci[k];
      if (j == i || std::fabs(A.v[k]) > theta * std::sqrt(aii * std::fabs(d[j]))) {
        C.ci.push_back(j);
        C.v.push_back(1.0);
      }
    }
    C.rs[i + 1] = (dlong)C.ci.size();
  }
  return C;
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
        if (c >= 0) t.push_back({r, c, Ae[((size_t)e * Np + a) * Np + b]});
      }
    }
  return fromTriplets(t, el.Ndofs, el.Ndofs);
}

amgSolver_t::amgSolver_t(elliptic_t& el, unsigned short xs[3]) : platform(el.platform) {
  LIBP_ABORT("AMG: the all-Neumann operator (singular) is not supported", el.allNeumann);
  build(ellipticAssembleCSR(el), xs);
}

amgSolver_t::amgSolver_t(platform_t& _platform, const csrHost_t& A, unsigned short xs[3]) : platform(_platform) {
  build(A, xs);
}

void amgSolver_t::build(const csrHost_t& A0, unsigned short xs[3]) {
  levels.clear();
  levels.emplace_back();
  levels[0].A = A0;
  std::vector<double> null(A0.n, 1.0);
  double theta = 0.08;  // halved on every level (parAlmondAMGSetup.cpp:86-119)
  for (;;) {
    level_t& L = levels.back();
    const csrHost_t& A = L.A;
    const std::vector<double> d = diagonal(A);
    L.invDiag.resize(A.n);
    for (dlong i = 0; i < A.n; ++i) L.invDiag[i] = (d[i] != 0.0) ? 1.0 / d[i] : 0.0;
    L.rho = rhoDinvA(A, L.invDiag, xs);
    L.lambda1 = L.rho;
    L.lambda0 = L.rho / 10.;
    if (A.n <= targetSize) break;
    // aggregates, tentative prolongator with normalised columns, smoothing
    const csrHost_t C = strongGraph(A, d, theta);
    dlong nc = 0;
    const std::vector<dlong> agg = formAggregates(C, xs, nc);
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
    // P = T - omega D^{-1} A T, omega = 4/3 / rho
    csrHost_t AT = multiply(A, T);
    const double omega = (4. / 3.) / L.rho;
    std::vector<triplet_t> t;
    for (dlong i = 0; i < A.n; ++i) {
      t.push_back({i, T.ci[i], T.v[i]});
      for (dlong k = AT.rs[i]; k < AT.rs[i + 1]; ++k) t.push_back({i, AT.ci[k], -omega * L.invDiag[i] * AT.v[k]});
    }
    L.P = fromTriplets(t, A.n, nc);
    L.R = transpose(L.P);
    csrHost_t Ac = multiply(L.R, multiply(A, L.P));
    const dlong n = A.n;
    levels.emplace_back();
    levels.back().A = std::move(Ac);
    null = std::move(nullc);
    theta *= 0.5;
    const dlong ncNew = levels.back().A.n;
    if (ncNew <= targetSize || n < 2 * ncNew) {
      level_t& Lc = levels.back();
      const std::vector<double> dc = diagonal(Lc.A);
      Lc.invDiag.resize(Lc.A.n);
      for (dlong i = 0; i < Lc.A.n; ++i) Lc.invDiag[i] = (dc[i] != 0.0) ? 1.0 / dc[i] : 0.0;
      break;
    }
  }
  // coarsest level: dense inverse by Gauss-Jordan with partial pivoting (parAlmondCoarseExact.cpp)
  const csrHost_t& Ac = levels.back().A;
  const dlong nc = Ac.n;
  std::vector<double> M((size_t)nc * nc, 0.0), I((size_t)nc * nc, 0.0);
  for (dlong i = 0; i < nc; ++i) {
    I[(size_t)i * nc + i] = 1.0;
    for (dlong k = Ac.rs[i]; k < Ac.rs[i + 1]; ++k) M[(size_t)i * nc + Ac.ci[k]] = Ac.v[k];
  }
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

// parCSR::smoothChebyshev (parAlmondAMGSmoother.cpp:68-160) with the vector
// updates fused as in the p-MG smoother
void amgSolver_t::smooth(level_t& L, const double* b, double* x, bool xIsZero) {
  const dlong N = L.A.n;
  const double theta = 0.5 * (L.lambda1 + L.lambda0), delta = 0.5 * (L.lambda1 - L.lambda0);
  const double sigma = theta / delta;
  double rho_n = 1. / sigma;
  void* s = platform.s();
  if (xIsZero) {
    LIBP_CALL(libp_cheb_start_f64(N, L.o_invDiag.ptr(), b, nullptr, 1.0 / theta, L.o_res.ptr(), L.o_d.ptr(), x, 1, s));
  } else {
    spmv(L.o_Ars, L.o_Aci, L.o_Av, N, L.lanesA, 1.0, x, 0.0, nullptr, L.o_Ad.ptr());
    LIBP_CALL(libp_cheb_start_f64(N, L.o_invDiag.ptr(), b, L.o_Ad.ptr(), 1.0 / theta, L.o_res.ptr(), L.o_d.ptr(), x,
                                  0, s));
  }
  for (int k = 0; k < chebDegree; ++k) {
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
  smooth(L, b, x, true);
  // res = b - A x ; bc = R res
  spmv(L.o_Ars, L.o_Aci, L.o_Av, L.A.n, L.lanesA, -1.0, x, 1.0, b, L.o_res.ptr());
  spmv(L.o_Rrs, L.o_Rci, L.o_Rv, L.R.n, L.lanesR, 1.0, L.o_res.ptr(), 0.0, nullptr, Lc.o_b.ptr());
  cycle(l + 1, Lc.o_b.ptr(), Lc.o_x.ptr());
  // x += P xc
  spmv(L.o_Prs, L.o_Pci, L.o_Pv, L.P.n, L.lanesP, 1.0, Lc.o_x.ptr(), 1.0, x, x);
  smooth(L, b, x, false);
}

void amgSolver_t::vcycle(const double* b, double* x) { cycle(0, b, x); }

}  // namespace libp
