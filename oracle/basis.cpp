// CPU oracle (test infrastructure only): 1-D basis, small dense linear algebra,
// and the two C-library random streams the reference depends on.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

/* Dense helpers. LAPACK (dgesv/dgetri/dgeev) stands behind these calls in
 * the reference (libs/linAlg/linAlgMatrixRightSolve.cpp:30,
 * linAlgMatrixInverse.cpp:48-52, linAlgMatrixEig.cpp:61,121). Here they are
 * plain partial-pivoting Gaussian elimination and a shifted QR iteration;
 * agreement with LAPACK is to rounding, which is why these are pinned by the
 * closed forms checked in the tests rather than bit-for-bit. */

// Solve X*A = B for X, with A n x n and B m x n (row-major). X is m x n.
static void rightSolve(int m, int n, const std::vector<double>& B,
                       const std::vector<double>& A, std::vector<double>& X) {
  // X*A = B  <=>  A^T X^T = B^T
  std::vector<double> M(n * n), R(n * m);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) M[i * n + j] = A[j * n + i];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) R[i * m + j] = B[j * n + i];
  std::vector<int> piv(n);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int i = c + 1; i < n; ++i)
      if (std::fabs(M[i * n + c]) > std::fabs(M[p * n + c])) p = i;
    if (M[p * n + c] == 0.0) throw std::runtime_error("singular matrix in rightSolve");
    if (p != c) {
      for (int j = 0; j < n; ++j) std::swap(M[c * n + j], M[p * n + j]);
      for (int j = 0; j < m; ++j) std::swap(R[c * m + j], R[p * m + j]);
    }
    for (int i = c + 1; i < n; ++i) {
      double f = M[i * n + c] / M[c * n + c];
      if (f == 0.0) continue;
      for (int j = c; j < n; ++j) M[i * n + j] -= f * M[c * n + j];
      for (int j = 0; j < m; ++j) R[i * m + j] -= f * R[c * m + j];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    for (int j = 0; j < m; ++j) {
      double v = R[i * m + j];
      for (int k = i + 1; k < n; ++k) v -= M[i * n + k] * R[k * m + j];
      R[i * m + j] = v / M[i * n + i];
    }
  }
  X.assign(m * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) X[j * n + i] = R[i * m + j];
}

static void matrixInverse(int n, std::vector<double>& A) {
  std::vector<double> I(n * n, 0.0), X;
  for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
  rightSolve(n, n, I, A, X);  // X*A = I
  A = X;
}

// Eigenvalues of a real symmetric tridiagonal matrix (diag d, off-diag e),
// implicit QL with Wilkinson shifts. Used in place of dgeev on the
// Golub-Welsch matrix of meshBasis1D.cpp:317-352.
static void tridiagEigs(std::vector<double> d, std::vector<double> e, std::vector<double>& out) {
  const int n = (int)d.size();
  e.push_back(0.0);
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-17 * dd) break;
      }
      if (m != l) {
        if (++iter > 200) throw std::runtime_error("tridiagEigs: no convergence");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  out = d;
}

// Eigenvalues of a general (upper Hessenberg) real matrix, column-major k x k as
// stored by ellipticPreconMultiGridLevel.cpp:432-452 (H[i + j*k]). Complex
// single-shift QR on the active block with deflation.
std::vector<std::complex<double>> hessenbergEigs(int n, const std::vector<double>& Hcm) {
  typedef std::complex<double> cd;
  std::vector<cd> A(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) A[i * n + j] = Hcm[i + j * n];
  double anorm = 0;
  for (auto& a : A) anorm = std::max(anorm, std::abs(a));
  if (anorm == 0) anorm = 1;
  std::vector<cd> ev;
  int hi = n - 1, iter = 0;
  std::vector<cd> cs(n), ss(n);
  while (hi >= 0) {
    if (hi == 0) { ev.push_back(A[0]); break; }
    int l = hi;
    while (l > 0) {
      double s = std::abs(A[(l - 1) * n + l - 1]) + std::abs(A[l * n + l]);
      if (s == 0) s = anorm;
      if (std::abs(A[l * n + l - 1]) <= 1e-16 * s) { A[l * n + l - 1] = 0; break; }
      --l;
    }
    if (l == hi) { ev.push_back(A[hi * n + hi]); --hi; iter = 0; continue; }
    cd a = A[(hi - 1) * n + hi - 1], b = A[(hi - 1) * n + hi], c = A[hi * n + hi - 1], d = A[hi * n + hi];
    cd tr = a + d, det = a * d - b * c;
    cd disc = std::sqrt(tr * tr / 4.0 - det);
    cd mu1 = tr / 2.0 + disc, mu2 = tr / 2.0 - disc;
    cd mu = (std::abs(mu1 - d) < std::abs(mu2 - d)) ? mu1 : mu2;
    if (iter > 0 && iter % 11 == 0) mu = d + std::abs(c);
    if (++iter > 1000) throw std::runtime_error("hessenbergEigs: no convergence");
    for (int i = l; i <= hi; ++i) A[i * n + i] -= mu;
    for (int i = l; i < hi; ++i) {
      cd x = A[i * n + i], y = A[(i + 1) * n + i];
      double r = std::sqrt(std::norm(x) + std::norm(y));
      cd cc = (r == 0) ? cd(1) : x / r, sc = (r == 0) ? cd(0) : y / r;
      cs[i] = cc; ss[i] = sc;
      for (int j = i; j <= hi; ++j) {
        cd u = A[i * n + j], v = A[(i + 1) * n + j];
        A[i * n + j] = std::conj(cc) * u + std::conj(sc) * v;
        A[(i + 1) * n + j] = -sc * u + cc * v;
      }
    }
    for (int i = l; i < hi; ++i) {
      cd cc = cs[i], sc = ss[i];
      int rmax = std::min(i + 2, hi);
      for (int rr = l; rr <= rmax; ++rr) {
        cd u = A[rr * n + i], v = A[rr * n + i + 1];
        A[rr * n + i] = u * cc + v * sc;
        A[rr * n + i + 1] = -u * std::conj(sc) + v * std::conj(cc);
      }
    }
    for (int i = l; i <= hi; ++i) A[i * n + i] += mu;
  }
  return ev;
}

/* ---------------- Jacobi polynomials: meshBasis1D.cpp:177-226 ---------------- */
double JacobiP(double a, double alpha, double beta, int N) {
  const double ax = a;
  std::vector<double> P(N + 1);
  const double gamma0 = std::pow(2, (alpha + beta + 1)) / (alpha + beta + 1) * std::tgamma(1 + alpha) *
                        std::tgamma(1 + beta) / std::tgamma(1 + alpha + beta);
  const double p0 = 1.0 / std::sqrt(gamma0);
  if (N == 0) return p0;
  P[0] = p0;
  const double gamma1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * gamma0;
  const double p1 = ((alpha + beta + 2) * ax / 2 + (alpha - beta) / 2) / std::sqrt(gamma1);
  if (N == 1) return p1;
  P[1] = p1;
  double aold = 2 / (2 + alpha + beta) * std::sqrt((alpha + 1.) * (beta + 1.) / (alpha + beta + 3.));
  for (int i = 1; i <= N - 1; ++i) {
    double h1 = 2. * i + alpha + beta;
    double anew = 2. / (h1 + 2.) *
                  std::sqrt((i + 1.) * (i + 1. + alpha + beta) * (i + 1 + alpha) * (i + 1 + beta) / (h1 + 1) / (h1 + 3));
    double bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2);
    P[i + 1] = 1. / anew * (-aold * P[i - 1] + (ax - bnew) * P[i]);
    aold = anew;
  }
  return P[N];
}

double GradJacobiP(double a, double alpha, double beta, int N) {
  double PNr = 0;
  if (N > 0) PNr = std::sqrt(N * (N + alpha + beta + 1.)) * JacobiP(a, alpha + 1.0, beta + 1.0, N - 1);
  return PNr;
}

/* Gauss-Jacobi nodes by the Golub-Welsch matrix (meshBasis1D.cpp:294-372).
 * Only the nodes are used by JacobiGLL, so the weights returned are the
 * closed-form Gauss-Jacobi weights for completeness. */
void JacobiGQ(double alpha, double beta, int N, std::vector<double>& x, std::vector<double>& w) {
  x.assign(N + 1, 0.0);
  w.assign(N + 1, 0.0);
  if (N == 0) {
    x[0] = (alpha - beta) / (alpha + beta + 2);
    w[0] = 2;
    return;
  }
  // Symmetric Golub-Welsch matrix J (+J' folded in), meshBasis1D.cpp:317-336.
  std::vector<double> d(N + 1), e(N);
  for (int n = 0; n <= N; ++n) {
    double h1 = 2 * n + alpha + beta;
    d[n] = -0.5 * (alpha * alpha - beta * beta) / ((h1 + 2) * h1) * 2;
    if (n < N)
      e[n] = (2. / (h1 + 2.)) * std::sqrt((n + 1) * (n + 1 + alpha + beta) * (n + 1 + alpha) * (n + 1 + beta) /
                                          ((h1 + 1) * (h1 + 3)));
  }
  double eps = 1;
  while (1 + eps > 1) eps = eps / 2.;
  if (alpha + beta < 10 * eps) d[0] = 0;
  tridiagEigs(d, e, x);
  std::sort(x.begin(), x.end());
}

/* GLL nodes and mass-lumped weights (meshBasis1D.cpp:246-289). */
void JacobiGLL(int N, std::vector<double>& x, std::vector<double>& w) {
  x.assign(N + 1, 0.0);
  w.assign(N + 1, 0.0);
  x[0] = -1.;
  x[N] = 1.;
  if (N > 1) {
    std::vector<double> xi, wi;
    JacobiGQ(1, 1, N - 2, xi, wi);
    for (int i = 0; i < N - 1; ++i) x[i + 1] = xi[i];
  }
  // V (Vandermonde1D, meshBasis1D.cpp:63-78) and MM = inv(V V') (:94-111)
  const int Np = N + 1;
  std::vector<double> V(Np * Np), MM(Np * Np);
  for (int n = 0; n < Np; ++n)
    for (int i = 0; i < Np; ++i) V[n * Np + i] = JacobiP(x[n], 0, 0, i);
  for (int n = 0; n < Np; ++n)
    for (int m = 0; m < Np; ++m) {
      double res = 0;
      for (int i = 0; i < Np; ++i) res += V[n * Np + i] * V[m * Np + i];
      MM[n * Np + m] = res;
    }
  matrixInverse(Np, MM);
  for (int n = 0; n < Np; ++n) {
    double res = 0;
    for (int m = 0; m < Np; ++m) res += MM[n * Np + m];
    w[n] = res;
  }
}

/* D = Vr / V  (meshBasis1D.cpp:114-136); D[i*Nq+m] = l'_m(r_i). */
void Dmatrix1D(int N, const std::vector<double>& r, std::vector<double>& D) {
  const int Np = N + 1;
  const int Nout = (int)r.size();
  std::vector<double> V(Np * Np), Vr(Nout * Np);
  for (int n = 0; n < Np; ++n)
    for (int i = 0; i < Np; ++i) V[n * Np + i] = JacobiP(r[n], 0, 0, i);
  for (int n = 0; n < Nout; ++n)
    for (int i = 0; i < Np; ++i) Vr[n * Np + i] = GradJacobiP(r[n], 0, 0, i);
  rightSolve(Nout, Np, Vr, V, D);
}

/* I = VOut / VIn (meshBasis1D.cpp:138-160). */
void InterpolationMatrix1D(int N, const std::vector<double>& rIn, const std::vector<double>& rOut,
                           std::vector<double>& I) {
  const int Np = N + 1;
  const int Nout = (int)rOut.size();
  std::vector<double> VIn(Np * Np), VOut(Nout * Np);
  for (int n = 0; n < Np; ++n)
    for (int i = 0; i < Np; ++i) VIn[n * Np + i] = JacobiP(rIn[n], 0, 0, i);
  for (int n = 0; n < Nout; ++n)
    for (int i = 0; i < Np; ++i) VOut[n * Np + i] = JacobiP(rOut[n], 0, 0, i);
  rightSolve(Nout, Np, VOut, VIn, I);
}

/* P (NqF x NqC, row-major): P[f*NqC + c] = l^c_c(r_f)  (meshBasis1D.cpp:162-170). */
void DegreeRaiseMatrix1D(int Nc, int Nf, std::vector<double>& P) {
  std::vector<double> rc, wc, rf, wf;
  JacobiGLL(Nc, rc, wc);
  JacobiGLL(Nf, rf, wf);
  InterpolationMatrix1D(Nc, rc, rf, P);
}

/* glibc random(): TYPE_3 additive feedback generator r[i] = r[i-3] + r[i-31],
 * seeded by the Park-Miller LCG (16807 mod 2^31-1), first 310 outputs
 * discarded, output r >> 1. rand() with no srand() call uses seed 1. The
 * reference calls rand() once per gathered node group in
 * libs/ogs/ogsSetup.cpp:264-275. */
void GlibcRand::seed(unsigned s) {
  int32_t r[34];
  r[0] = (int32_t)(s == 0 ? 1 : s);
  for (int i = 1; i < 31; ++i) {
    int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
    int64_t word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    r[i] = (int32_t)word;
  }
  for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
  for (int i = 0; i < 34; ++i) ring[i] = (uint32_t)r[i];
  pos = 0;  // ring holds r[k..k+33]; pos points at r[k]
  for (int i = 34; i < 344; ++i) next();
}

int GlibcRand::next() {
  // ring[pos] = r[i-34]; r[i-31] = ring[(pos+3)%34], r[i-3] = ring[(pos+31)%34]
  uint32_t v = ring[(pos + 3) % 34] + ring[(pos + 31) % 34];
  ring[pos] = v;
  pos = (pos + 1) % 34;
  return (int)(v >> 1);
}

}  // namespace oracle
