// Aggregation algebraic multigrid for the p-multigrid base level.
//
// Reference: libs/parAlmond (parAlmondAMGSetup.cpp:35-144 setup loop,
// parAlmondStrongGraph.cpp:39-222 symmetric or Ruge-Stuben strength,
// parAlmondFormAggregates.cpp MIS-2 aggregation, parAlmondTentativeProlongator.cpp,
// parAlmondSmoothPrologator.cpp P = (I - 4/3 / rho D^{-1} A) T or the unsmoothed
// P = T (parAlmondCoarsenLevel.cpp:54-67), Ac = P^T (A P) or the Galerkin product
// of parAlmondGalerkinProd.cpp, parCSR::rhoDinvA (parAlmondparCSR.cpp:364-460)
// 10-step Arnoldi, parAlmondAMGLevel.cpp:76-84 Chebyshev bounds [rho/10, rho] or
// damped Jacobi 4/3 / rho, parAlmondCoarseExact.cpp dense inverse at <= 1000 rows,
// parAlmondVcycle.cpp the cycle). The matrix is the assembled operator of one
// elliptic level (ellipticBuildOperatorMatrixContinuous.cpp:697-880: element
// entries below 1e-7 in magnitude dropped), formed here by applying the element
// operator to unit vectors. Setup runs on the host in FP64; the V-cycle runs on
// the device (CSR SpMV, Chebyshev or damped Jacobi, dense GEMV).
#pragma once

#include <complex>
#include <vector>

#include "core.hpp"

namespace libp {

class elliptic_t;

// eigenvalues of a k x k upper Hessenberg matrix, column-major H[i + j k] (multigrid.cpp)
std::vector<std::complex<double>> hessenbergEigenvalues(int n, const std::vector<double>& H);

struct csrHost_t {
  dlong n = 0, m = 0;         // rows, columns
  std::vector<dlong> rs, ci;  // row starts (n + 1), column ids (ascending in each row)
  std::vector<double> v;
  dlong nnz() const { return rs.empty() ? 0 : rs[n]; }
};

/* Assembled operator of an elliptic level on its gathered (unmasked) dofs:
 * the element matrices come from the element operator applied to unit
 * vectors (libp_ax_hex_local_f64), summed over the copies of each node. */
csrHost_t ellipticAssembleCSR(elliptic_t& el);

class amgSolver_t {
 public:
  struct level_t {
    csrHost_t A, P, R;  // P: n x nc to the next level, R = P^T (empty on the coarsest level)
    std::vector<double> invDiag;
    std::vector<dlong> aggregate;  // aggregate (coarse row) of each row (empty on the coarsest level)
    double rho = 0, lambda = 0, lambda0 = 0, lambda1 = 0;
    int lanesA = 1, lanesP = 1, lanesR = 1;
    deviceMemory<dlong> o_Ars, o_Aci, o_Prs, o_Pci, o_Rrs, o_Rci;
    deviceMemory<double> o_Av, o_Pv, o_Rv, o_invDiag;
    deviceMemory<double> o_b, o_x, o_res, o_d, o_Ad;
  };
  // PARALMOND settings (parAlmondSettings.cpp:36-59)
  struct options_t {
    bool smoothedAggregation = true;  // PARALMOND AGGREGATION SMOOTHED | UNSMOOTHED
    bool rugeStuben = false;          // PARALMOND STRENGTH SYMMETRIC | RUGESTUBEN
    bool chebyshev = true;            // PARALMOND SMOOTHER CHEBYSHEV | DAMPEDJACOBI
    int chebDegree = 2;               // PARALMOND CHEBYSHEV DEGREE
  };

  platform_t platform;
  options_t opt;
  std::vector<level_t> levels;  // levels[0] is the operator itself; the last one is solved exactly
  std::vector<double> null;     // null vector of the coarsest level
  deviceMemory<double> o_coarseInv;
  int targetSize = 1000;

  // Options from the settings; aborts on PARALMOND CYCLE values other than VCYCLE.
  static options_t optionsFrom(const settings_t& settings);

  // xs: erand48 state for the Arnoldi starts and aggregation weights (the reference
  // draws both from drand48). nullSpace adds penalty * null null^T to the coarsest
  // matrix before it is inverted (the all-Neumann operator, parAlmondCoarseExact.cpp).
  amgSolver_t(elliptic_t& el, unsigned short xs[3]);
  amgSolver_t(platform_t& platform, const csrHost_t& A, unsigned short xs[3], const options_t& opt,
              bool nullSpace = false, double nullSpacePenalty = 0.0);

  // x = V(b): one V-cycle from a zero guess; device FP64 vectors of the finest level
  void vcycle(const double* b, double* x);
  // one smoothing of level l (x updated in place; xIsZero ignores x's contents)
  void smooth(size_t l, const double* b, double* x, bool xIsZero);

 private:
  void build(const csrHost_t& A, unsigned short xs[3], bool nullSpace, double nullSpacePenalty);
  void cycle(size_t l, const double* b, double* x);
  void spmv(const deviceMemory<dlong>& rs, const deviceMemory<dlong>& ci, const deviceMemory<double>& v, dlong n,
            int lanes, double alpha, const double* x, double beta, const double* z, double* y);
};

}  // namespace libp
