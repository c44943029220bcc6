// (F)PCG on the device (libs/linearSolver/linearSolverPCG.cpp:35-171).
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "solver.hpp"

namespace libp {
namespace LinearSolver {

pcg::pcg(dlong _N, dlong _Nhalo, platform_t& _platform, settings_t& _settings, comm_t _comm)
    : linearSolverBase_t(_N, _Nhalo, _platform, _settings, _comm) {
  flexible = settings.compareSetting("LINEAR SOLVER", "FPCG");
  fuseJacobi = !settings.compareSetting("PCG JACOBI FUSED", "FALSE");
  if (settings.hasSetting("LINEAR SOLVER CHECK EVERY")) settings.getSetting("LINEAR SOLVER CHECK EVERY", checkEvery);
  const dlong Ntotal = N + Nhalo;
  o_p = platform.malloc<dfloat>(Ntotal);
  o_z = platform.malloc<dfloat>(Ntotal);
  o_Ax = platform.malloc<dfloat>(Ntotal);
  o_Ap = platform.malloc<dfloat>(Ntotal);
  o_s = platform.malloc<dfloat>(LIBP_PCG_NSCALARS);
  // zero work vectors (linearSolverPCG.cpp:45-50 fills them with 0)
  LIBP_CALL(libp_set_f64(Ntotal, 0.0, o_p.ptr(), platform.s()));
  LIBP_CALL(libp_set_f64(Ntotal, 0.0, o_z.ptr(), platform.s()));
  LIBP_CALL(libp_set_f64(Ntotal, 0.0, o_Ax.ptr(), platform.s()));
  LIBP_CALL(libp_set_f64(Ntotal, 0.0, o_Ap.ptr(), platform.s()));
}

/* One PCG iteration (linearSolverPCG.cpp:107-171). With a Jacobi
 * preconditioner, a plain iteration and an operator that has the fused form
 * (operator_t::OperatorJacobiPCG), an iteration is three passes: p = z + beta
 * p; the Ax kernel with p.Ap and alpha; the gather fused with the x, r, z
 * update, r.r, r.z and beta (Ap is never stored).
 * Otherwise, with [PCG JACOBI FUSED]
 * the operator's gather also forms p.Ap whatever the preconditioner, and with
 * a Jacobi preconditioner and a plain (not flexible) iteration the launches
 * fuse further:
 *   p = z + beta p; Ap = A p with p.Ap and alpha in the gather
 *   (operator_t::OperatorPAP); x, r updated with z = invDiag.*r, r.r, r.z and
 *   beta for the next iteration (libp_pcg_jacobi_update_f64).
 * The first iteration forms z and r.z of the initial residual first. The
 * arithmetic of every entry is that of the separate kernels; only the
 * partitions of the reductions differ. */
void pcg::iteration(operator_t& linearOperator, operator_t& precon, diagonalPrecon_t* jac, int it,
                    deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r) {
  void* st = platform.s();
  void* ws = platform.o_reduceWS.ptr();
  double* s = o_s.ptr();
  const bool fused = jac && !flexible && fuseJacobi;
  if (fused) {
    if (it == 0)
      LIBP_CALL(libp_pcg_jacobi_rdotz_f64(N, jac->invDiagonal().ptr(), o_r.ptr(), o_z.ptr(), o_Ap.ptr(), 0, s, ws,
                                          st));
    LIBP_CALL(libp_pcg_update_p_f64(N, it == 0, o_z.ptr(), o_p.ptr(), s, st));
    if (linearOperator.OperatorJacobiPCG(o_p, o_x, o_r, o_z, jac->invDiagonal(), s)) return;
    if (!linearOperator.OperatorPAP(o_p, o_Ap, s, ws)) {
      linearOperator.Operator(o_p, o_Ap);
      LIBP_CALL(libp_pcg_pap_f64(N, o_p.ptr(), o_Ap.ptr(), s, ws, st));
    }
    LIBP_CALL(libp_pcg_jacobi_update_f64(N, jac->invDiagonal().ptr(), o_p.ptr(), o_Ap.ptr(), o_x.ptr(), o_r.ptr(),
                                         o_z.ptr(), s, ws, st));
    return;
  }
  if (jac) {
    LIBP_CALL(libp_pcg_jacobi_rdotz_f64(N, jac->invDiagonal().ptr(), o_r.ptr(), o_z.ptr(), o_Ap.ptr(), flexible, s,
                                        ws, st));
  } else if (!fuseJacobi || !precon.OperatorRdotz(o_r, o_z, o_Ap, flexible, s, ws)) {
    precon.Operator(o_r, o_z);
    LIBP_CALL(libp_pcg_rdotz_f64(N, o_r.ptr(), o_z.ptr(), o_Ap.ptr(), flexible, s, ws, st));
  }
  LIBP_CALL(libp_pcg_update_p_f64(N, it == 0, o_z.ptr(), o_p.ptr(), s, st));
  if (!fuseJacobi || !linearOperator.OperatorPAP(o_p, o_Ap, s, ws)) {
    linearOperator.Operator(o_p, o_Ap);
    LIBP_CALL(libp_pcg_pap_f64(N, o_p.ptr(), o_Ap.ptr(), s, ws, st));
  }
  LIBP_CALL(libp_pcg_update_f64(N, o_p.ptr(), o_Ap.ptr(), o_x.ptr(), o_r.ptr(), s, ws, st));
}

int pcg::Solve(operator_t& linearOperator, operator_t& precon, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r,
               const dfloat tol, const int MAXIT, const int verbose) {
  void* st = platform.s();
  void* ws = platform.o_reduceWS.ptr();
  double* s = o_s.ptr();
  double* tmp = platform.o_scalars.ptr();
  stats = solveStats_t();
  LIBP_CALL(libp_set_f64(LIBP_PCG_NSCALARS, 0.0, s, st));

  dfloat TOL = 0.0;
  const bool rhsNorm = settings.compareSetting("LINEAR SOLVER STOPPING CRITERION", "ABS/REL-RHS-2NORM");
  if (rhsNorm) {
    LIBP_CALL(libp_norm2sq_f64(N, o_r.ptr(), tmp, ws, st));
    const double nb2 = platform.readScalars(tmp, 1)[0];
    TOL = std::max(tol * tol * nb2, tol * tol);
  }

  initialResidual(linearOperator, o_x, o_Ax, o_r);
  LIBP_CALL(libp_norm2sq_f64(N, o_r.ptr(), tmp, ws, st));
  double rdotr0 = platform.readScalars(tmp, 1)[0];
  if (!rhsNorm) TOL = pureRelative ? tol * tol * rdotr0 : std::max(tol * tol * rdotr0, tol * tol);
  stats.rdotr0 = rdotr0;
  if (recordHistory) stats.resHistory.push_back(std::sqrt(rdotr0));
  if (verbose) printf("PCG: initial res norm %12.12f \n", std::sqrt(rdotr0));

  precon.solveStarting();
  // a diagonal preconditioner, also behind the precon_t wrapper that elliptic_t holds
  operator_t* inner = &precon;
  if (auto* w = dynamic_cast<precon_t*>(&precon))
    if (w->get()) inner = w->get();
  diagonalPrecon_t* jac = dynamic_cast<diagonalPrecon_t*>(inner);
  if (jac && jac->needsZeroMean()) jac = nullptr;

  // Chunked convergence checks need device-side scalars only: a Jacobi (or no)
  // preconditioner, no per-iteration callbacks, no history or printing.
  if (checkEvery > 1 && jac && !recordHistory && !verbose) {
    // chunked: the kernels stop themselves once rdotr <= TOL (device STOP flag)
    int iter = 0;
    if (rdotr0 != 0.0) {
      const double tolv = TOL > 0 ? TOL : 1e-300;
      LIBP_CALL(libp_set_f64(1, tolv, s + LIBP_PCG_TOL, st));
      while (iter < MAXIT) {
        const int chunk = std::min(checkEvery, MAXIT - iter);
        for (int c = 0; c < chunk; ++c) {
          const int it = iter + c;
          iteration(linearOperator, precon, jac, it, o_x, o_r);
        }
        const double* h = platform.readScalars(s, LIBP_PCG_NSCALARS);
        rdotr0 = h[LIBP_PCG_RDOTR];
        iter = (int)h[LIBP_PCG_ITERS];
        if (h[LIBP_PCG_STOP] != 0.0) break;
      }
    }
    stats.iterations = iter;
    stats.rdotrFinal = rdotr0;
    return iter;
  }

  int iter;
  for (iter = 0; iter < MAXIT; ++iter) {
    if (((iter == 0) && (rdotr0 == 0.0)) || ((iter > 0) && (rdotr0 <= TOL))) break;
    iteration(linearOperator, precon, jac, iter, o_x, o_r);
    const double* h = platform.readScalars(s, LIBP_PCG_NSCALARS);
    rdotr0 = h[LIBP_PCG_RDOTR];
    precon.iterationDone(iter, rdotr0);
    if (recordHistory) stats.resHistory.push_back(std::sqrt(std::max(rdotr0, 0.0)));
    if (verbose) {
      if (rdotr0 < 0) printf("WARNING CG: rdotr = %17.15lf\n", rdotr0);
      printf("CG: it %d, r norm %12.12le, alpha = %le \n", iter + 1, std::sqrt(rdotr0), h[LIBP_PCG_ALPHA]);
    }
  }
  stats.iterations = iter;
  stats.rdotrFinal = rdotr0;
  return iter;
}

}  // namespace LinearSolver
}  // namespace libp
