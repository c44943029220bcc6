// operator_t / precon_t / linearSolver_t and the (F)PCG solver.
// Interfaces follow include/operator.hpp:35-40, include/precon.hpp:36-69 and
// include/linearSolver.hpp:41-96 so elliptic_t, the preconditioners and
// callers written against them keep their shape.
#pragma once

#include "core.hpp"

namespace libp {

class operator_t {
 public:
  virtual void Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) = 0;
  // Aq = A q together with the PCG scalars scalars[PAP] = q.Aq and
  // scalars[ALPHA] = scalars[RDOTZ1] / q.Aq on the device (libp_pcg_pap_f64).
  // Returns false when the operator has no fused form; the solver then runs
  // Operator and the separate reduction.
  virtual bool OperatorPAP(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_Aq, double* scalars, void* ws) {
    (void)o_q;
    (void)o_Aq;
    (void)scalars;
    (void)ws;
    return false;
  }
  // One Jacobi-PCG iteration after the p update, with p.Ap formed by the Ax
  // kernel (libp_pcg_ax_hex_pap_f64) and the operator's gather fused into the
  // update (libp_pcg_gather_jacobi_update_f64: x += alpha p, r -= alpha Ap,
  // z = invDiag.*r, r.r, r.z and beta; Ap never stored). Returns false when
  // the operator has no such form.
  virtual bool OperatorJacobiPCG(deviceMemory<dfloat>& o_p, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r,
                                 deviceMemory<dfloat>& o_z, const deviceMemory<dfloat>& o_invDiag, double* scalars) {
    (void)o_p;
    (void)o_x;
    (void)o_r;
    (void)o_z;
    (void)o_invDiag;
    (void)scalars;
    return false;
  }
  // For a preconditioner: z = M r together with the PCG scalars of
  // libp_pcg_rdotz_f64 (r.z, and z.Ap when flexible) in one pass over z.
  // Returns false when there is no fused form.
  virtual bool OperatorRdotz(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_z, deviceMemory<dfloat>& o_Ap,
                             int flexible, double* scalars, void* ws) {
    (void)o_r;
    (void)o_z;
    (void)o_Ap;
    (void)flexible;
    (void)scalars;
    (void)ws;
    return false;
  }
  // Solver notifications (no-ops by default): a solve starts; iteration `it`
  // ended with residual norm^2 rdotr (already on the host).
  virtual void solveStarting() {}
  virtual void iterationDone(int it, double rdotr) {
    (void)it;
    (void)rdotr;
  }
  virtual ~operator_t() {}
};

// A preconditioner that is a diagonal scaling (Jacobi) exposes its inverse
// diagonal so the solver can fuse z = invDiag.*r into the r.z reduction.
class diagonalPrecon_t : public operator_t {
 public:
  virtual const deviceMemory<dfloat>& invDiagonal() const = 0;
  virtual bool needsZeroMean() const { return false; }
};

class precon_t : public operator_t {
  std::shared_ptr<operator_t> precon;

 public:
  template <class Precon, class... Args>
  void Setup(Args&&... args) {
    precon = std::make_shared<Precon>(std::forward<Args>(args)...);
  }
  void Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) override { precon->Operator(o_r, o_Mr); }
  bool OperatorRdotz(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_z, deviceMemory<dfloat>& o_Ap, int flexible,
                     double* scalars, void* ws) override {
    return precon && precon->OperatorRdotz(o_r, o_z, o_Ap, flexible, scalars, ws);
  }
  void solveStarting() override {
    if (precon) precon->solveStarting();
  }
  void iterationDone(int it, double rdotr) override {
    if (precon) precon->iterationDone(it, rdotr);
  }
  operator_t* get() const { return precon.get(); }
  bool isInitialized() const { return (bool)precon; }
};

class IdentityPrecon : public operator_t {
  dlong N;
  platform_t platform;

 public:
  IdentityPrecon(dlong _N, platform_t& _platform) : N(_N), platform(_platform) {}
  void Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) override {
    LIBP_HIP(hipMemcpyAsync(o_Mr.ptr(), o_r.ptr(), sizeof(dfloat) * N, hipMemcpyDeviceToDevice, platform.stream));
  }
};

struct solveStats_t {
  int iterations = 0;
  double rdotr0 = 0, rdotrFinal = 0;
  std::vector<double> resHistory;  // |r| after each iteration (index 0: initial)
};

class linearSolverBase_t {
 public:
  platform_t platform;
  settings_t settings;
  comm_t comm;
  dlong N = 0, Nhalo = 0;
  solveStats_t stats;
  bool pureRelative = false;  // stop at |r| <= tol |r0| (no absolute floor)
  int checkEvery = 1;         // host convergence test every k iterations (the device stops itself)
  bool recordHistory = false;
  // Set by a caller whose initial guess is exactly zero (elliptic_t::Run gathers
  // a zero E-vector into x): the next Solve then takes r = rhs without forming
  // A x, which is zero, so r - A x is rhs bit for bit (up to the sign of zero
  // entries). The Solve clears it.
  bool zeroInitialGuess = false;
  // r -= A x, the initial residual of linearSolverPCG.cpp:88-91, skipped for a
  // zero initial guess
  void initialResidual(operator_t& A, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_Ax,
                       deviceMemory<dfloat>& o_r) {
    const bool zero = zeroInitialGuess;
    zeroInitialGuess = false;
    if (zero) return;
    A.Operator(o_x, o_Ax);
    LIBP_CALL(libp_axpy_f64(N, -1.0, o_Ax.ptr(), 1.0, o_r.ptr(), platform.s()));
  }

  linearSolverBase_t(dlong _N, dlong _Nhalo, platform_t& _platform, settings_t& _settings, comm_t _comm)
      : platform(_platform), settings(_settings), comm(_comm), N(_N), Nhalo(_Nhalo) {}
  virtual int Solve(operator_t& linearOperator, operator_t& precon, deviceMemory<dfloat>& o_x,
                    deviceMemory<dfloat>& o_rhs, const dfloat tol, const int MAXIT, const int verbose) = 0;
  virtual ~linearSolverBase_t() {}
};

class linearSolver_t {
  std::shared_ptr<linearSolverBase_t> ls;

 public:
  template <class Solver>
  void Setup(dlong N, dlong Nhalo, platform_t& platform, settings_t& settings, comm_t comm) {
    ls = std::make_shared<Solver>(N, Nhalo, platform, settings, comm);
  }
  int Solve(operator_t& A, operator_t& M, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_rhs, const dfloat tol,
            const int MAXIT, const int verbose) {
    return ls->Solve(A, M, o_x, o_rhs, tol, MAXIT, verbose);
  }
  linearSolverBase_t* get() const { return ls.get(); }
};

namespace LinearSolver {

/* (F)PCG with device-resident scalars (libs/linearSolver/linearSolverPCG.cpp).
 * alpha and beta are formed by the reduction kernels on the device; the host
 * reads r.r once per iteration for the stopping test. Flexible when the
 * LINEAR SOLVER setting contains "FPCG" (linearSolverPCG.cpp:43). */
class pcg : public linearSolverBase_t {
  bool flexible = false;
  bool fuseJacobi = true;  // [PCG JACOBI FUSED]: p.Ap in the gather; fused Jacobi-PCG iterations
  deviceMemory<dfloat> o_p, o_z, o_Ax, o_Ap;
  deviceMemory<dfloat> o_s;  // LIBP_PCG_NSCALARS device scalars
  void iteration(operator_t& linearOperator, operator_t& precon, diagonalPrecon_t* jac, int it,
                 deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r);

 public:
  pcg(dlong N, dlong Nhalo, platform_t& platform, settings_t& settings, comm_t comm);
  int Solve(operator_t& linearOperator, operator_t& precon, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r,
            const dfloat tol, const int MAXIT, const int verbose) override;
  void setFlexible(bool f) { flexible = f; }
};

/* Non-blocking (pipelined) PCG (libs/linearSolver/linearSolverNBPCG.cpp):
 * two fused update/multi-dot kernels per iteration, alpha/beta/gamma on the
 * device; the preconditioner must be linear (z is updated by recurrence). */
class nbpcg : public linearSolverBase_t {
  deviceMemory<dfloat> o_p, o_s, o_S, o_z, o_Z, o_Ax, o_sc;

 public:
  nbpcg(dlong N, dlong Nhalo, platform_t& platform, settings_t& settings, comm_t comm);
  int Solve(operator_t& linearOperator, operator_t& precon, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r,
            const dfloat tol, const int MAXIT, const int verbose) override;
};

/* Non-blocking flexible PCG, Sanan's pipelined variant
 * (libs/linearSolver/linearSolverNBFPCG.cpp), scalars on the device. */
class nbfpcg : public linearSolverBase_t {
  deviceMemory<dfloat> o_u, o_p, o_w, o_n, o_m, o_s, o_z, o_q, o_Ax, o_sc;

 public:
  nbfpcg(dlong N, dlong Nhalo, platform_t& platform, settings_t& settings, comm_t comm);
  int Solve(operator_t& linearOperator, operator_t& precon, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r,
            const dfloat tol, const int MAXIT, const int verbose) override;
};

}  // namespace LinearSolver
}  // namespace libp
