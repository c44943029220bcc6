// Non-blocking PCG and FPCG (libs/linearSolver/linearSolverNBPCG.cpp:71-160,
// linearSolverNBFPCG.cpp:71-181) with the fused update kernels of nbpcg.hip.
#include <cmath>
#include <cstdio>

#include "solver.hpp"

namespace libp {
namespace LinearSolver {

namespace {
enum { G = 0, G1 = 1, ALPHA = 2, BETA = 3, DELTA = 4, RR = 5, ZZ = 6 };          // NBPCG scalars
enum { FG = 0, FDELTA = 1, FETA = 2, FALPHA = 3, FBETA = 4, FRR = 5 };             // NBFPCG scalars
const int NSC = 8;
}  // namespace

nbpcg::nbpcg(dlong _N, dlong _Nhalo, platform_t& _platform, settings_t& _settings, comm_t _comm)
    : linearSolverBase_t(_N, _Nhalo, _platform, _settings, _comm) {
  const dlong Nt = N + Nhalo;
  o_p = platform.malloc<dfloat>(Nt);
  o_s = platform.malloc<dfloat>(Nt);
  o_S = platform.malloc<dfloat>(Nt);
  o_z = platform.malloc<dfloat>(Nt);
  o_Z = platform.malloc<dfloat>(Nt);
  o_Ax = platform.malloc<dfloat>(Nt);
  o_sc = platform.malloc<dfloat>(NSC);
  for (auto* v : {&o_p, &o_s, &o_S, &o_z, &o_Z, &o_Ax}) LIBP_CALL(libp_set_f64(Nt, 0.0, v->ptr(), platform.s()));
}

int nbpcg::Solve(operator_t& A, operator_t& M, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r, const dfloat tol,
                 const int MAXIT, const int verbose) {
  void* st = platform.s();
  void* ws = platform.o_reduceWS.ptr();
  double* sc = o_sc.ptr();
  stats = solveStats_t();
  LIBP_CALL(libp_set_f64(NSC, 0.0, sc, st));
  initialResidual(A, o_x, o_Ax, o_r);
  M.solveStarting();
  M.Operator(o_r, o_z);
  LIBP_CALL(libp_nbpcg_update2_f64(N, 1, o_s.ptr(), o_S.ptr(), o_r.ptr(), o_z.ptr(), sc, ws, st));
  A.Operator(o_z, o_Z);
  double rdotr = platform.readScalars(sc, NSC)[RR];
  const double TOL = pureRelative ? tol * tol * rdotr : std::max(tol * tol * rdotr, tol * tol);
  stats.rdotr0 = rdotr;
  if (recordHistory) stats.resHistory.push_back(std::sqrt(rdotr));
  if (verbose) printf("NBPCG: initial res norm %12.12f \n", std::sqrt(rdotr));
  int iter;
  for (iter = 0; iter < MAXIT; ++iter) {
    if (rdotr <= TOL) break;
    LIBP_CALL(libp_nbpcg_update1_f64(N, o_z.ptr(), o_Z.ptr(), o_p.ptr(), o_s.ptr(), sc, ws, st));
    M.Operator(o_s, o_S);
    LIBP_CALL(libp_nbpcg_update2_f64(N, 0, o_s.ptr(), o_S.ptr(), o_r.ptr(), o_z.ptr(), sc, ws, st));
    LIBP_CALL(libp_axpy_dev_f64(N, sc + ALPHA, o_p.ptr(), o_x.ptr(), st));
    A.Operator(o_z, o_Z);
    const double* h = platform.readScalars(sc, NSC);
    rdotr = h[RR];
    M.iterationDone(iter, rdotr);
    if (recordHistory) stats.resHistory.push_back(std::sqrt(std::max(rdotr, 0.0)));
    if (verbose) printf("NBPCG: it %d, r norm %12.12le, gamma = %le zdotz = %le \n", iter + 1, std::sqrt(rdotr), h[G], h[ZZ]);
  }
  stats.iterations = iter;
  stats.rdotrFinal = rdotr;
  return iter;
}

nbfpcg::nbfpcg(dlong _N, dlong _Nhalo, platform_t& _platform, settings_t& _settings, comm_t _comm)
    : linearSolverBase_t(_N, _Nhalo, _platform, _settings, _comm) {
  const dlong Nt = N + Nhalo;
  for (auto* v : {&o_u, &o_p, &o_w, &o_n, &o_m, &o_s, &o_z, &o_q, &o_Ax}) {
    *v = platform.malloc<dfloat>(Nt);
    LIBP_CALL(libp_set_f64(Nt, 0.0, v->ptr(), platform.s()));
  }
  o_sc = platform.malloc<dfloat>(NSC);
}

int nbfpcg::Solve(operator_t& A, operator_t& M, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r,
                  const dfloat tol, const int MAXIT, const int verbose) {
  void* st = platform.s();
  void* ws = platform.o_reduceWS.ptr();
  double* sc = o_sc.ptr();
  const size_t bytes = (size_t)N * sizeof(double);
  stats = solveStats_t();
  LIBP_CALL(libp_set_f64(NSC, 0.0, sc, st));
  initialResidual(A, o_x, o_Ax, o_r);
  M.solveStarting();
  M.Operator(o_r, o_u);
  LIBP_HIP(hipMemcpyAsync(o_p.ptr(), o_u.ptr(), bytes, hipMemcpyDeviceToDevice, platform.stream));
  A.Operator(o_p, o_w);
  LIBP_CALL(libp_nbfpcg_update0_f64(N, o_u.ptr(), o_r.ptr(), o_w.ptr(), sc, ws, st));
  M.Operator(o_w, o_m);
  A.Operator(o_m, o_n);
  LIBP_HIP(hipMemcpyAsync(o_s.ptr(), o_w.ptr(), bytes, hipMemcpyDeviceToDevice, platform.stream));
  LIBP_HIP(hipMemcpyAsync(o_q.ptr(), o_m.ptr(), bytes, hipMemcpyDeviceToDevice, platform.stream));
  LIBP_HIP(hipMemcpyAsync(o_z.ptr(), o_n.ptr(), bytes, hipMemcpyDeviceToDevice, platform.stream));
  double rdotr = platform.readScalars(sc, NSC)[FRR];
  const double TOL = pureRelative ? tol * tol * rdotr : std::max(tol * tol * rdotr, tol * tol);
  stats.rdotr0 = rdotr;
  if (recordHistory) stats.resHistory.push_back(std::sqrt(rdotr));
  if (verbose) printf("NBFPCG: initial res norm %12.12f \n", std::sqrt(rdotr));
  int iter;
  for (iter = 0; iter < MAXIT; ++iter) {
    if (rdotr <= TOL) break;
    LIBP_CALL(libp_nbfpcg_update1_f64(N, o_p.ptr(), o_s.ptr(), o_q.ptr(), o_z.ptr(), o_x.ptr(), o_r.ptr(), o_u.ptr(),
                                      o_w.ptr(), sc, ws, st));
    LIBP_CALL(libp_zaxpy_f64(N, 1.0, o_w.ptr(), -1.0, o_r.ptr(), o_n.ptr(), st));
    M.Operator(o_n, o_m);
    LIBP_CALL(libp_axpy_f64(N, 1.0, o_u.ptr(), 1.0, o_m.ptr(), st));
    A.Operator(o_m, o_n);
    LIBP_CALL(libp_nbfpcg_update2_f64(N, o_u.ptr(), o_w.ptr(), o_m.ptr(), o_n.ptr(), o_p.ptr(), o_s.ptr(), o_q.ptr(),
                                      o_z.ptr(), sc, st));
    const double* h = platform.readScalars(sc, NSC);
    rdotr = h[FRR];
    M.iterationDone(iter, rdotr);
    if (recordHistory) stats.resHistory.push_back(std::sqrt(std::max(rdotr, 0.0)));
    if (verbose)
      printf("NBFPCG: it %d, r norm %12.12le, alpha = %12.12le beta = %le rdotr = %le gamma = %le delta = %le, eta = %le \n",
             iter + 1, std::sqrt(rdotr), h[FALPHA], h[FBETA], rdotr, h[FG], h[FDELTA], h[FETA]);
  }
  stats.iterations = iter;
  stats.rdotrFinal = rdotr;
  return iter;
}

}  // namespace LinearSolver
}  // namespace libp
