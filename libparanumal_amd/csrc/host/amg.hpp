// Smoothed-aggregation algebraic multigrid for the p-multigrid base level.
//
// Reference: libs/parAlmond (parAlmondAMGSetup.cpp:35-144 setup loop,
// parAlmondStrongGraph.cpp:147-230 symmetric strength, parAlmondFormAggregates.cpp
// MIS-2 aggregation, parAlmondTentativeProlongator.cpp, parAlmondSmoothPrologator.cpp
// P = (I - 4/3 / rho D^{-1} A) T, parAlmondCoarsenLevel.cpp Ac = P^T (A P),
// parAlmondAMGLevel.cpp:76-84 Chebyshev bounds [rho/10, rho], parAlmondCoarseExact.cpp
// dense inverse at <= 1000 rows, parAlmondVcycle.cpp the cycle). The matrix is
// the assembled operator of one elliptic level (ellipticBuildOperatorMatrixContinuous
// restated by applying the element operator to unit vectors). Setup runs on the
// host in FP64; the V-cycle runs on the device (CSR SpMV, Chebyshev, dense GEMV).
#pragma once

#include <vector>

#include "core.hpp"

namespace libp {

class elliptic_t;

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
    double rho = 0, lambda0 = 0, lambda1 = 0;
    int lanesA = 1, lanesP = 1, lanesR = 1;
    deviceMemory<dlong> o_Ars, o_Aci, o_Prs, o_Pci, o_Rrs, o_Rci;
    deviceMemory<double> o_Av, o_Pv, o_Rv, o_invDiag;
    deviceMemory<double> o_b, o_x, o_res, o_d, o_Ad;
  };

  platform_t platform;
  std::vector<level_t> levels;  // levels[0] is the operator itself; the last one is solved exactly
  deviceMemory<double> o_coarseInv;
  int chebDegree = 2;
  int targetSize = 1000;

  // xs: erand48 state for the aggregation weights (the reference draws them with drand48)
  amgSolver_t(elliptic_t& el, unsigned short xs[3]);
  amgSolver_t(platform_t& platform, const csrHost_t& A, unsigned short xs[3]);

  // x = V(b): one V-cycle from a zero guess; device FP64 vectors of the finest level
  void vcycle(const double* b, double* x);

 private:
  void build(const csrHost_t& A, unsigned short xs[3]);
  void cycle(size_t l, const double* b, double* x);
  void smooth(level_t& L, const double* b, double* x, bool xIsZero);
  void spmv(const deviceMemory<dlong>& rs, const deviceMemory<dlong>& ci, const deviceMemory<double>& v, dlong n,
            int lanes, double alpha, const double* x, double beta, const double* z, double* y);
};

}  // namespace libp
