// elliptic_t (C0 hex) and its preconditioners for one MI355X GPU.
// Shape follows solvers/elliptic/elliptic.hpp:56-170 and ellipticPrecon.hpp:34-185.
#pragma once

#include "mesh.hpp"
#include "ogs.hpp"
#include "solver.hpp"

namespace libp {

struct ellipticRunResult_t {
  double norm = 0;        // "Solution norm" (ellipticRun.cpp:238-247)
  int iterations = 0;
  double solveSeconds = 0;
  double rFinal = 0, r0 = 0;
  double maxError = NAN;  // max nodal |u_h - u| for the sine data set
};

/* Maps of the split operator (libp_ax_hex_split_f64 + libp_ogs_gather_face_f64):
 * the gathered rows that have element-boundary copies and, in the slot layout,
 * each boundary node's place in the row-ordered list of copies (ascending local
 * order within a row, as in the ogs CSR), or in the element layout each row's
 * copies as indices e * Nface + face index; the face E-vector and the p.Ap
 * partials of the fused PCG step.
 * Rows of element-interior nodes have one copy and are written by the Ax kernel. */
struct splitGather_t {
  int Nface = 0;
  dlong NfaceRows = 0;
  bool slotLayout = true;  // AqFace in row-ordered copy slots (o_slots) instead of element order (o_cols)
  deviceMemory<dlong> o_rows, o_rowStarts, o_cols, o_slots;
  deviceMemory<dfloat> o_AqFace;
  deviceMemory<double> o_papPart;
  int nAxPartials = 0;
};

class elliptic_t : public operator_t {
 public:
  platform_t platform;
  mesh_t mesh;
  comm_t comm;
  settings_t settings;
  std::shared_ptr<randStream_t> rng;  // process-wide rand() stream of the ogs setups

  int Nfields = 1;
  int disc_c0 = 1;
  dfloat lambda = 0;
  int NBCTypes = 3;
  memory<int> BCType;
  int allNeumann = 0;
  dfloat allNeumannPenalty = 1, allNeumannScale = 0;

  memory<int> EToB, mapB;
  deviceMemory<int> o_mapB;
  dlong Nmasked = 0;
  memory<dlong> maskIds;
  memory<hlong> maskedGlobalIds;
  ogs::ogs_t ogsMasked;
  memory<dfloat> weightG;
  deviceMemory<dfloat> o_weightG;
  deviceMemory<float> o_weightG32;
  memory<dlong> GlobalToLocal;
  deviceMemory<dlong> o_GlobalToLocal;
  deviceMemory<dfloat> o_AqL;
  deviceMemory<float> o_AqL32;
  std::shared_ptr<splitGather_t> split;  // built on first use ([OPERATOR SPLIT GATHER] TRUE, N <= 14)
  bool splitTried = false;
  dlong Ndofs = 0, Nhalo = 0;
  precon_t precon;

  elliptic_t() {}
  elliptic_t(platform_t& platform, mesh_t& mesh, settings_t& settings, dfloat lambda, const int NBCTypes,
             const memory<int> BCType, std::shared_ptr<randStream_t> rng = nullptr) {
    Setup(platform, mesh, settings, lambda, NBCTypes, BCType, rng);
  }
  // ellipticSetup.cpp:30-180
  void Setup(platform_t& platform, mesh_t& mesh, settings_t& settings, dfloat lambda, const int NBCTypes,
             const memory<int> BCType, std::shared_ptr<randStream_t> rng = nullptr);
  void BoundarySetup();  // ellipticBoundarySetup.cpp:30-140
  void SetupPrecon();    // ellipticSetup.cpp:166-179

  // ellipticOperator.cpp:29-107 (single launch over all local elements)
  void Operator(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_Aq) override;
  // the partial Ax into the level's own E-vector buffer (returned), for callers
  // that fuse the gather with what follows it
  double* PartialAxLocal(deviceMemory<dfloat>& o_q);
  float* PartialAxLocalF32(deviceMemory<float>& o_q);
  // res = r - A q, the subtraction fused into the operator's gather (FP64 and FP32)
  void OperatorResidual(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_res);
  void OperatorResidualF32(deviceMemory<float>& o_q, deviceMemory<float>& o_r, deviceMemory<float>& o_res);
  // the operator's gather fused with the PCG p.Ap (libp_pcg_gather_pap_f64)
  bool OperatorPAP(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_Aq, double* scalars, void* ws) override;
  // the Jacobi-PCG iteration with p.Ap in the Ax and the gather fused into the update
  // ([PCG FUSED OPERATOR] TRUE, N <= 15, two-pass operator; N = 15 on the MFMA kernel)
  bool OperatorJacobiPCG(deviceMemory<dfloat>& o_p, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r,
                         deviceMemory<dfloat>& o_z, const deviceMemory<dfloat>& o_invDiag, double* scalars) override;
  deviceMemory<double> o_axPartials, o_gatherPartials;
  void PartialAx(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_AqL);
  // the split operator's maps, or null when the two-pass form is used
  splitGather_t* SplitGather();
  // FP32 operator used by the p-MG smoother (no reference counterpart)
  void OperatorF32(deviceMemory<float>& o_q, deviceMemory<float>& o_Aq);

  void BuildOperatorDiagonal(deviceMemory<dfloat>& o_diagA);  // ellipticBuildOperatorDiagonal.cpp:29-80
  void ZeroMean(deviceMemory<dfloat>& o_q);                   // ellipticZeroMean.cpp
  elliptic_t SetupNewDegree(mesh_t& meshC);                    // ellipticSetupNewDegree.cpp:30-142

  // ellipticRun.cpp:30-248 and ellipticSolve.cpp:29-39
  ellipticRunResult_t Run(int verbose = -1);
  int Solve(linearSolver_t& linearSolver, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r, const dfloat tol,
            const int MAXIT, const int verbose);
  // Data set of the DATA FILE setting: kind (0 sine, 1 Fourier) and device coefficients
  int dataKind() const;
  deviceMemory<dfloat> dataCoefficients();
  deviceMemory<dfloat> dataCoefficients(long seed);
  // gathered right-hand side (forcing + Dirichlet lift) of the data with the given seed
  void BuildRhs(long seed, deviceMemory<dfloat>& o_r);

  deviceMemory<dfloat> o_xL_last;  // solution E-vector of the last Run
  // ellipticPlotFields.cpp: the E-vector interpolated to (N+4)^3 equispaced nodes per
  // element, six tetrahedra per sub-cube, ASCII VTU (one file per rank)
  void PlotFields(const deviceMemory<dfloat>& o_qL, const std::string& fileName);
};

class JacobiPrecon : public diagonalPrecon_t {
 public:
  elliptic_t& elliptic;
  deviceMemory<dfloat> o_invDiagA;
  explicit JacobiPrecon(elliptic_t& elliptic);
  void Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) override;
  const deviceMemory<dfloat>& invDiagonal() const override { return o_invDiagA; }
  bool needsZeroMean() const override { return elliptic.allNeumann != 0; }
};

class MultiGridPreconBase;
class MultiGridPrecon : public operator_t {
 public:
  elliptic_t& elliptic;
  std::shared_ptr<MultiGridPreconBase> impl;
  explicit MultiGridPrecon(elliptic_t& elliptic);
  void Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) override;
  bool OperatorRdotz(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_z, deviceMemory<dfloat>& o_Ap, int flexible,
                     double* scalars, void* ws) override;
  void solveStarting() override;
  void iterationDone(int it, double rdotr) override;
};

class learnedSmoother_t;
// Fine-level learned smoother of an FP32 p-MG preconditioner (null otherwise).
learnedSmoother_t* multigridLearned(MultiGridPrecon& mg);
int multigridCoarseIterations(MultiGridPrecon& mg);
// FP32 inverse diagonal and operator of the fine p-MG level (FP32 preconditioner only)
deviceMemory<float> multigridFineInvDiag32(MultiGridPrecon& mg);
void multigridApplyFineOperator32(MultiGridPrecon& mg, deviceMemory<float>& x, deviceMemory<float>& y);
// Level access for tests: 32 or 64 (the level precision), the gathered length of
// level l, one operation of level l (0 A, 1 coarsen to l+1, 2 prolongate from l+1
// (out += P in), 3 smooth (in = rhs, out = x in place, flag = x is zero), 4 the
// inverse diagonal) on device vectors of that precision, and its Chebyshev
// interval {lambda1, lambda0}.
class amgSolver_t;
// the AMG base level of a p-MG preconditioner ([MULTIGRID COARSE SOLVER] AMG), or null
amgSolver_t* multigridBaseAmg(MultiGridPrecon& mg);
int multigridPrecisionBits(MultiGridPrecon& mg);
long multigridLevelNdofs(MultiGridPrecon& mg, int l);
void multigridLevelOp(MultiGridPrecon& mg, int l, int op, const void* in, void* out, int flag);
void multigridLevelLambda(MultiGridPrecon& mg, int l, double* lambda);

/* Training of the learned fine smoother (no reference counterpart). Pairs
 * (r, eh) come from FPCG solves preconditioned by this p-MG with the Chebyshev
 * smoother, one per seed of `seeds` (Fourier forcing of that seed): r is each
 * smoother input residual (outer residuals, and the residuals after the
 * coarse correction), scaled to unit norm; eh = A^{-1} r - P A_1^{-1} R r is
 * the error left after an exact coarse-grid correction. `dagger` further
 * rounds add pairs from solves that use the network being trained. Returns
 * the mean loss of each block of `logEvery` steps. */
struct learnedTrainOptions {
  std::vector<long> seeds{101, 102, 103, 104};
  std::vector<long> daggerSeeds{201, 202};
  int steps = 3000;
  int dagger = 1;
  int daggerSteps = 1500;
  float lr = 1e-3f;
  double targetTol = 1e-9;
  unsigned long initSeed = 1;
  int logEvery = 250;
  int maxRecordPerSolve = 24;
  // linear architecture (least squares on white-noise error samples)
  int linearSamples = 16;
  double ridge = 1e-4;
  double linearCoarseTol = 1e-6;
};
std::vector<double> multigridTrainLearned(MultiGridPrecon& mg, const learnedTrainOptions& opt, int verbose);

// Degree schedule of MultiGridPrecon (ellipticPreconMultiGrid.cpp:45-131) plus CUSTOM.
std::vector<int> multigridDegrees(const settings_t& settings, int N);

}  // namespace libp
