// CPU oracle (test infrastructure only): a driver that runs every restated
// stage on small inputs, built with -fsanitize=address,undefined by
// `make -C oracle sanitize` so that memory and undefined-behaviour errors in
// the oracle surface without a Python interpreter in the process. Exit status
// 0 when every run finished with a finite solution norm.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "oracle.hpp"

using namespace oracle;

static Settings settingsFrom(const std::string& text) {
  Settings s = Settings::defaults();
  Settings user;
  user.parseText(text);
  for (auto& k : user.order) s.set(k, user.get(k));
  return s;
}

static std::string box(int n, int N, const char* extra) {
  char buf[1024];
  std::snprintf(buf, sizeof(buf),
                "[MESH FILE]\nBOX\n[MESH DIMENSION]\n3\n[ELEMENT TYPE]\n12\n[BOX NX]\n%d\n[BOX NY]\n%d\n"
                "[BOX NZ]\n%d\n[BOX DIMX]\n2\n[BOX DIMY]\n2\n[BOX DIMZ]\n2\n[BOX BOUNDARY FLAG]\n1\n"
                "[POLYNOMIAL DEGREE]\n%d\n[BOX CURVED MAP AMPLITUDE]\n0.05\n[DATA FILE]\ndata/ellipticSine3D.h\n%s",
                n, n + 1, n, N, extra);
  return buf;
}

int main() {
  int failures = 0;
  const char* runs[][2] = {
      {"pcg-none", "[LAMBDA]\n1\n[PRECONDITIONER]\nNONE\n"},
      {"pcg-jacobi", "[LAMBDA]\n0\n[PRECONDITIONER]\nJACOBI\n"},
      {"nbpcg-jacobi", "[LAMBDA]\n1\n[LINEAR SOLVER]\nNBPCG\n[PRECONDITIONER]\nJACOBI\n"},
      {"fpcg-pmg-jacobi-pcg-base",
       "[LAMBDA]\n1\n[LINEAR SOLVER]\nFPCG\n[PRECONDITIONER]\nMULTIGRID\n[MULTIGRID COARSENING]\nCUSTOM\n"
       "[MULTIGRID DEGREES]\n4,2,1\n[LINEAR SOLVER STOPPING CRITERION]\nRELATIVE\n"},
      {"fpcg-pmg-amg-base",
       "[LAMBDA]\n1\n[LINEAR SOLVER]\nFPCG\n[PRECONDITIONER]\nMULTIGRID\n[MULTIGRID COARSE SOLVER]\nAMG\n"
       "[PARALMOND AGGREGATION]\nUNSMOOTHED\n"},
      {"nbfpcg-pmg-damped-jacobi",
       "[LAMBDA]\n1\n[LINEAR SOLVER]\nNBFPCG\n[PRECONDITIONER]\nMULTIGRID\n[MULTIGRID SMOOTHER]\nDAMPEDJACOBI\n"},
  };
  for (auto& r : runs) {
    try {
      RunResult res = elliptic_run(settingsFrom(box(3, 4, r[1])), 0);
      const bool ok = std::isfinite(res.norm) && res.iters > 0;
      std::printf("%-28s iterations %4d norm %.15f %s\n", r[0], res.iters, res.norm, ok ? "ok" : "FAILED");
      failures += !ok;
    } catch (std::exception& e) {
      std::printf("%-28s exception: %s\n", r[0], e.what());
      ++failures;
    }
  }
  // AMG with several levels on the assembled N=1 operator
  try {
    Settings s = settingsFrom(box(12, 1, "[LAMBDA]\n0\n[PARALMOND SMOOTHER]\nDAMPEDJACOBI\n"));
    GlibcRand rng;
    Mesh mesh;
    mesh.Setup(s);
    Elliptic el;
    el.Setup(mesh, s, 0.0, rng);
    Drand48 d;
    d.seed(0);
    AmgO amg;
    amg.Setup(BuildOperatorMatrixContinuousHex3D(el), s, d, false, 0.0);
    std::vector<double> b(el.Ndofs, 1.0), x(el.Ndofs, 0.0);
    amg.vcycle(0, b.data(), x.data());
    double nx = 0;
    for (double v : x) nx += v * v;
    const bool ok = std::isfinite(nx) && amg.levels.size() >= 2;
    std::printf("%-28s levels %zu |x|^2 %.6e %s\n", "amg-vcycle", amg.levels.size(), nx, ok ? "ok" : "FAILED");
    failures += !ok;
    // learned smoother oracle: traces, forward, loss and gradient of a small MLP
    Settings s7 = settingsFrom(box(2, 3, "[LAMBDA]\n0\n"));
    Mesh m7;
    m7.Setup(s7);
    Elliptic e7;
    GlibcRand rng7;
    e7.Setup(m7, s7, 0.0, rng7);
    std::vector<int> tr;
    LearnedTraceIds(e7, tr);
    const int Np = m7.Np, Nfp = m7.Nq * m7.Nq, nIn = Np + 6 * Nfp, H = 8;
    const size_t nth = (size_t)H * nIn + H + H * H + H + (size_t)Np * H + Np + 1;
    std::vector<double> th(nth), r(e7.Ndofs), eh(e7.Ndofs), invD(e7.Ndofs, 0.5), g(nth), c(e7.Ndofs);
    for (size_t i = 0; i < nth; ++i) th[i] = 0.01 * std::sin(0.7 * (double)i);
    for (dlong i = 0; i < e7.Ndofs; ++i) {
      r[i] = std::cos(0.3 * i);
      eh[i] = std::sin(0.2 * i);
    }
    const double J = LearnedLossGrad(e7, tr, H, th.data(), r.data(), eh.data(), 1.0, invD.data(), true, g.data(),
                                     c.data());
    const bool ok2 = std::isfinite(J);
    std::printf("%-28s loss %.6e %s\n", "learned-loss-grad", J, ok2 ? "ok" : "FAILED");
    failures += !ok2;
  } catch (std::exception& e) {
    std::printf("amg/learned exception: %s\n", e.what());
    ++failures;
  }
  std::printf("%s\n", failures ? "SANITIZE CHECK FAILED" : "SANITIZE CHECK OK");
  return failures ? 1 : 0;
}
