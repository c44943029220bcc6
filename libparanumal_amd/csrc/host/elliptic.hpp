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
  void PartialAx(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_AqL);
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

  deviceMemory<dfloat> o_xL_last;  // solution E-vector of the last Run
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
};

// Degree schedule of MultiGridPrecon (ellipticPreconMultiGrid.cpp:45-131) plus CUSTOM.
std::vector<int> multigridDegrees(const settings_t& settings, int N);

}  // namespace libp
