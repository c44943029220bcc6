/*
 * CPU oracle for the libParanumal C0 hex elliptic path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain C++ restatement of the reference's
 * Serial-mode (OCCA "Serial") behaviour, used as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg. The product
 * (libparanumal_amd) never links, loads or calls anything in this directory.
 *
 * Each function cites the reference file:line it restates (paths relative to the
 * libParanumal tree). Single rank, single process: every MPI call of the
 * reference is a size-1 no-op here and is dropped.
 */
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace oracle {

typedef double dfloat;     // include/types.h:31-45 (dfloat=double)
typedef int dlong;         // include/types.h:53-60 (dlong=int)
typedef long long hlong;   // include/types.h:62-69 (hlong=long long)

/* ------------------------------------------------------------------ */
/* settings (.rc) — libs/core/settings.cpp:134-214, :73-75             */
/* ------------------------------------------------------------------ */
struct Settings {
  std::map<std::string, std::string> kv;
  std::vector<std::string> order;
  void set(const std::string& name, const std::string& val);
  bool has(const std::string& name) const;
  std::string get(const std::string& name) const;          // aborts if missing
  bool compare(const std::string& name, const std::string& token) const; // substring
  int geti(const std::string& name) const;
  double getd(const std::string& name) const;
  void parseText(const std::string& text);                  // settings.cpp:134-214
  static Settings defaults();                                // elliptic/mesh/platform/parAlmond defaults
};

/* ------------------------------------------------------------------ */
/* 1-D basis — libs/mesh/meshBasis1D.cpp                              */
/* ------------------------------------------------------------------ */
double JacobiP(double a, double alpha, double beta, int N);      // meshBasis1D.cpp:186-216
double GradJacobiP(double a, double alpha, double beta, int N);  // meshBasis1D.cpp:218-226
void JacobiGQ(double alpha, double beta, int N, std::vector<double>& x, std::vector<double>& w); // :294-372
void JacobiGLL(int N, std::vector<double>& x, std::vector<double>& w);  // :246-289
void Dmatrix1D(int N, const std::vector<double>& r, std::vector<double>& D);   // :114-136
void InterpolationMatrix1D(int N, const std::vector<double>& rIn,
                           const std::vector<double>& rOut, std::vector<double>& I); // :138-160
void DegreeRaiseMatrix1D(int Nc, int Nf, std::vector<double>& P);  // :162-170

/* glibc rand() (TYPE_3 additive feedback, default seed 1) and POSIX drand48 */
struct GlibcRand {
  uint32_t ring[34];
  int pos = 0;
  GlibcRand() { seed(1); }
  void seed(unsigned s);
  int next();
};
struct Drand48 {
  uint64_t x;
  void seed(long s) { x = ((uint64_t)(uint32_t)s << 16) | 0x330E; }
  double next() {
    x = (0x5DEECE66DULL * x + 0xB) & ((1ULL << 48) - 1);
    return (double)x / (double)(1ULL << 48);
  }
};

/* ------------------------------------------------------------------ */
/* mesh — libs/mesh/                                                   */
/* ------------------------------------------------------------------ */
struct Mesh {
  int N = 0, Nq = 0, Np = 0, Nfp = 0;
  static const int Nverts = 8, Nfaces = 6, NfaceVertices = 4;
  dlong Nelements = 0;
  hlong Nnodes = 0;
  dlong nx = 0, ny = 0, nz = 0;
  int boundaryFlag = 1;
  double DIMX = 2, DIMY = 2, DIMZ = 2;
  // curved map (no reference counterpart): x += amp*sin(2 pi y)*sin(2 pi z)
  double curveAmp = 0.0;

  std::vector<hlong> EToV;
  std::vector<double> EX, EY, EZ;
  hlong NboundaryFaces = 0;
  std::vector<hlong> boundaryInfo;

  std::vector<double> r, s, t, gllz, gllw, D;
  std::vector<int> faceNodes;
  std::vector<int> faceVertices;

  std::vector<dlong> EToE;
  std::vector<int> EToF, EToB;
  std::vector<dlong> vmapM, vmapP;
  std::vector<hlong> globalIds;
  std::vector<int> mapB;

  std::vector<double> x, y, z;
  std::vector<double> ggeo, wJ;   // ggeo [e][6][Np], wJ [e][Np]

  std::vector<dlong> localGatherElementList;

  void SetupBox(const Settings& s);             // meshSetupBoxHex3D.cpp:31-232
  void ReferenceNodes(int N);                   // meshReferenceNodesHex3D.cpp:31-74
  void Connect();                               // meshConnect.cpp:51-…
  void ConnectBoundary();                       // meshConnectBoundary.cpp:54-156
  void ConnectFaceNodes();                      // meshConnectFaceNodes.cpp:32-101
  void ConnectNodes();                          // meshConnectNodes.cpp:32-117
  void PhysicalNodes();                         // meshPhysicalNodesHex3D.cpp:31-112 (+ curved map)
  void GeometricFactors();                      // meshGeometricFactorsHex3D.cpp:31-193
  void GatherScatterSetup();                    // meshGatherScatterSetup.cpp:32-130 (1 rank)
  void Setup(const Settings& s);                // meshSetup.cpp:31-102
  Mesh SetupNewDegree(int Nf) const;            // meshSetupNewDegree.cpp:32-64
};

/* ------------------------------------------------------------------ */
/* ogs — libs/ogs/ogsSetup.cpp, ogsOperator.cpp                       */
/* ------------------------------------------------------------------ */
struct OgsOperator {
  dlong NrowsN = 0, NrowsT = 0, Ncols = 0;
  std::vector<dlong> rowStartsN, rowStartsT, colIdsN, colIdsT;
  dlong NrowBlocksN = 0, NrowBlocksT = 0;
  std::vector<dlong> blockRowStartsN, blockRowStartsT;
  static const int gatherNodesPerBlock = 512;   // include/ogs/ogsOperator.hpp:122-123
  void setupRowBlocks();                        // ogsOperator.cpp:495-560
  template <typename T> void GatherAdd(T* gv, const T* v, bool trans, bool parallel = false) const;  // :64-113
  template <typename T> void Scatter(T* v, const T* gv, bool trans) const;     // :225-266
};

struct Ogs {
  dlong N = 0;
  dlong Ngather = 0, NlocalT = 0, NlocalP = 0;
  OgsOperator gatherLocal;
  // ids is modified in place (owner sign written back) like ogsSetup.cpp:137-147
  void Setup(dlong N, std::vector<hlong>& ids, bool unique, GlibcRand& rng);  // ogsSetup.cpp:69-190
  void SetupGlobalToLocalMapping(std::vector<dlong>& G2L) const;            // ogsSetup.cpp:867-886
};

/* ------------------------------------------------------------------ */
/* elliptic — solvers/elliptic/src/                                    */
/* ------------------------------------------------------------------ */
struct Elliptic;

struct MGLevelO {
  Elliptic* el = nullptr;       // this level
  Elliptic* elC = nullptr;      // next coarser level
  int Nc = 0;
  std::vector<double> P;        // NqF x NqC, row-major
  std::vector<double> invDiagA;
  int stype = 1;                // 1 = Chebyshev, 0 = damped Jacobi
  int ChebyshevIterations = 2;
  double lambda0 = 0, lambda1 = 0;
};

struct Elliptic {
  Mesh mesh;
  Settings settings;
  double lambda = 0;
  int allNeumann = 0;
  double allNeumannPenalty = 1.0, allNeumannScale = 0;
  std::vector<int> EToB, mapB, maskIds;
  dlong Nmasked = 0;
  std::vector<hlong> maskedGlobalIds;
  Ogs ogsMasked;
  std::vector<double> weight, weightG;
  std::vector<dlong> GlobalToLocal;
  dlong Ndofs = 0;
  bool parallel = false;  // Operator runs the element and row loops under OpenMP (CPU baseline timing)

  void Setup(const Mesh& mesh, const Settings& s, double lambda, GlibcRand& rng);
  void BoundarySetup(GlibcRand& rng);            // ellipticBoundarySetup.cpp:30-140
  // partial Ax over element list (ellipticAxHex3D.okl:156-295)
  void PartialAx(const double* q, double* AqL, bool parallel = false) const;
  void Operator(const double* q, double* Aq, std::vector<double>& scratchL) const; // ellipticOperator.cpp:29-107
  void BuildOperatorDiagonal(std::vector<double>& diagA) const; // ellipticBuildOperatorDiagonal.cpp:29-80,998-1057
  void ZeroMean(double* q) const;                // ellipticZeroMean.cpp
  void Rhs(std::vector<double>& rL) const;       // ellipticRhsHex3D.okl:28-51 + ellipticRhsBCHex3D.okl
  void AddBC(std::vector<double>& xL) const;     // ellipticAddBCHex3D.okl:27-47
};

/* BLAS-1 restatements with the OKL block structure (libs/linAlg/okl) */
double linalg_innerProd(dlong N, const double* x, const double* y);  // linAlgInnerProd.okl + linAlg.cpp:192-207
double linalg_norm2(dlong N, const double* x);                        // linAlgNorm2.okl + linAlg.cpp:176-190
double linalg_sum(dlong N, const double* x);

/* ------------------------------------------------------------------ */
/* solver                                                              */
/* ------------------------------------------------------------------ */
struct OpFn {
  virtual void apply(const double* x, double* y) = 0;
  virtual ~OpFn() {}
};

struct PCGResult {
  int iters = 0;
  double rdotr0 = 0, rdotrFinal = 0;
  std::vector<double> resHistory;
};

// libs/linearSolver/linearSolverPCG.cpp:67-171
PCGResult pcg_solve(dlong N, OpFn& A, OpFn& M, double* x, double* r, double tol, int maxit,
                    bool flexible, bool stopRhsNorm, bool pureRelative = false);

/* parAlmond AMG (amg.cpp), one rank */
struct CsrO {
  dlong n = 0, m = 0;
  std::vector<dlong> rs, ci;  // rows sorted, columns ascending in each row
  std::vector<double> v;
  dlong nnz() const { return rs.empty() ? 0 : rs[n]; }
  void spmv(double alpha, const double* x, double beta, double* y) const;  // y = alpha A x + beta y
};
struct AmgLevelO {
  CsrO A, P, R;                      // P, R empty on the coarsest level
  std::vector<double> diagA, diagInv;
  double rho = 0;                    // parCSR::rhoDinvA
  std::vector<dlong> Crs, Cci;       // strong graph
  std::vector<hlong> FineToCoarse;   // aggregate of each row
  std::vector<double> weights;       // MIS-2 weights: drand48 draw + strong column count
  std::vector<int> states;           // MIS-2 result: 1 root, -1 removed
  int stype = 1;                     // 1 Chebyshev, 0 damped Jacobi
  int ChebyshevIterations = 2;
  double lambda = 0, lambda0 = 0, lambda1 = 0;
};
struct AmgO {
  std::vector<AmgLevelO> levels;
  std::vector<double> null;          // null vector of the coarsest level
  std::vector<double> coarseInv;     // dense inverse of the coarsest operator, row-major
  void Setup(const CsrO& A, const Settings& s, Drand48& drng, bool nullSpace, double nullSpacePenalty);
  void smooth(int k, const double* b, double* x, bool xIsZero) const;
  void vcycle(int k, const double* b, double* x) const;
};
// ellipticBuildOperatorMatrixContinuous.cpp:697-880 (hex, one rank)
CsrO BuildOperatorMatrixContinuousHex3D(const Elliptic& el);
std::vector<double> denseInverse(int n, std::vector<double> M);

struct MultiGrid {
  std::unique_ptr<AmgO> amg;   // base level: parAlmond AMG when MULTIGRID COARSE SOLVER is AMG
  std::vector<std::unique_ptr<Elliptic>> levels;   // levels[0] = fine (shallow copy of the fine elliptic)
  std::vector<MGLevelO> mg;
  std::vector<std::vector<double>> invDiagCoarse;
  int coarseMaxIt = 500;
  double coarseTol = 1e-2;
  bool fp32 = false;   // emulate FP32 smoothing / transfers by rounding
  void Setup(Elliptic& fine, GlibcRand& rng, Drand48& drng, const std::vector<int>& degrees);
  void vcycle(int k, std::vector<double>& rhs, std::vector<double>& x);
  void apply(const double* r, double* z);
  std::vector<std::vector<double>> rhsL, xL;
};

/* p-multigrid level operations (ellipticPreconMultiGridLevel.cpp), exposed so
 * that tests can pin each one against closed forms on small dense systems. */
double maxEigSmoothAx(const Elliptic& el, const std::vector<double>& invDiagA, Drand48& drng);  // :393-473
void smoothChebyshev(const MGLevelO& L, const double* r, double* X, bool xIsZero);             // :154-206
void smoothJacobi(const MGLevelO& L, const double* r, double* X, bool xIsZero);                // :135-152
void coarsen(const MGLevelO& L, const double* X, double* Rx);                                  // :41-82
void prolongate(const MGLevelO& L, const double* Xc, double* Px);                              // :84-125

/* End-to-end run (ellipticRun.cpp:30-248). Returns the solution norm. */
struct RunResult {
  double norm = 0;
  int iters = 0;
  dlong Ngather = 0;
  std::vector<double> xL;
};
RunResult elliptic_run(const Settings& s, int verbose = 0);

/* learned smoother oracle (learned.cpp); no reference counterpart */
double bf16r(double x);
void LearnedTraceIds(const Elliptic& el, std::vector<int>& trace);
void LearnedLinearGram(const Elliptic& el, const std::vector<int>& trace, const double* r, const double* eh,
                       const double* invD, bool bf, double* G, double* B);
double LearnedLossGrad(const Elliptic& el, const std::vector<int>& trace, int H, const double* th, const double* r,
                       const double* eh, double nrm, const double* invD, bool bf, double* grad, double* cOut);
void LearnedForward(const Elliptic& el, const std::vector<int>& trace, int H, const double* th, const double* r,
                    const double* invD, bool bf, double* X, double* scale, double* Y, double* c);

}  // namespace oracle
