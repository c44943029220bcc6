// Learned per-element smoother for the fine p-MG level (no reference counterpart).
#pragma once

#include <functional>
#include <string>

#include "core.hpp"

namespace libp {

class elliptic_t;

/* A per-element MLP (nIn = Np + 6 Nfp -> 256 -> 256 -> Np, ReLU, BF16 weights
 * with FP32 masters, FP32 accumulation on MFMA) maps the Jacobi-scaled
 * residual of an element and of the six neighbour layers next to its faces to
 * an element error correction. Element corrections are averaged over the
 * copies of each node and a learned multiple omega of the Jacobi correction
 * is added:
 *   c(r) = avg_e[s_e N(z_e / s_e)] + omega D^{-1} r,  z = D^{-1} r,
 * with s_e the RMS of the element's input. Each application takes the
 * energy-optimal step x += alpha c, alpha = c.res / c.Ac, which never raises
 * the A-norm error; a non-finite or non-positive c.Ac rejects the step.
 *
 * Parameters are one FP32 vector theta (layout below), written to disk with a
 * header and an FNV-1a hash. Training minimises the energy norm
 *   J = (eh - c(r)).A(eh - c(r)) / eh.A eh
 * over residual / target pairs, where eh is the part of the error A^{-1} r that
 * the coarse-grid correction leaves (computed by the multigrid trainer).
 *
 * theta: W1 [H][nIn], b1 [H], W2 [H][H], b2 [H], W3 [Np][H], b3 [Np], omega.
 *
 * ARCH_LINEAR replaces the three layers by one bias-free linear layer
 * y_e = W x_e (theta: W [Np][nIn], omega) and the energy-optimal step by the
 * fixed step x += c: the smoother is then a fixed linear map of r, an
 * element-overlapping Schwarz-type local solve whose local operator is
 * learned, and the V-cycle a fixed linear preconditioner. It is fitted by
 * least squares: for white-noise errors v (r = A v, eh = v - P A_1^{-1} R r) it
 * minimises sum_e |W x_e - eh_e / s_e|^2 + ridge, from Gram matrices formed by
 * the backward-weight GEMM. `sweeps` applications make one smoothing. */
class learnedSmoother_t {
 public:
  typedef std::function<void(deviceMemory<float>&, deviceMemory<float>&)> op32_t;
  // (x, r, res): res = r - A x
  typedef std::function<void(deviceMemory<float>&, deviceMemory<float>&, deviceMemory<float>&)> res32_t;

  enum { ARCH_MLP = 0, ARCH_LINEAR = 1 };

  elliptic_t& elliptic;
  platform_t platform;
  int arch = ARCH_MLP;
  int sweeps = 1;  // ARCH_LINEAR: applications per smoothing
  int Np = 0, Nfp = 0, nIn = 0, H = 256, nOut = 0;
  dlong E = 0, Ndofs = 0;
  size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, oOm = 0, nParams = 0;

  bool active = false;    // weights loaded and accepted
  bool fallback = false;  // switched to Chebyshev at run time
  std::string status = "no weights";
  int rejectedSteps = 0;

  deviceMemory<int> o_trace;    // [E][6 Nfp] gathered ids of the neighbour traces (-1: none)
  deviceMemory<float> o_invLen;  // [Ndofs] 1 / multiplicity
  deviceMemory<float> o_theta, o_grad, o_m, o_v;
  deviceMemory<unsigned short> o_thetaBf;
  // LINEAR at N=7: the layer fused with its input (libp_learned_linear_layer_f32) reads
  // the weights in the MFMA-packed order, repacked whenever the BF16 weights change
  bool fusedLayer = false;
  deviceMemory<unsigned short> o_Wpacked;
  deviceMemory<unsigned short> o_X, o_H1, o_H2, o_dY, o_dH1, o_dH2, o_T;
  deviceMemory<float> o_scale, o_Y, o_c, o_Ac, o_res, o_gc, o_d, o_Ad;
  deviceMemory<double> o_sc;  // [0..3] step statistics, [4] last loss, [5] running loss sum
  deviceMemory<char> o_ws, o_rws;
  int adamStep = 0;
  float omegaHost = 0.f;  // omega as last set from the host (the linear fit never changes it)
  double monitorRejected = 0;
  std::vector<double> monitorHist;

  explicit learnedSmoother_t(elliptic_t& el, int arch = ARCH_MLP);

  // trace map from the box structure; checked against the global node ids
  void setupMaps();
  void initRandom(unsigned long seed);
  // false (and status set) when the file is missing, malformed, of another
  // shape, fails its hash or holds non-finite values
  bool load(const std::string& path);
  void save(const std::string& path) const;
  std::vector<float> parameters() const;
  void setParameters(const std::vector<float>& th);

  // c = c(r) in o_c (features, three layers, assembly)
  // the correction c(r) into o_c, or into out (added to it when accumulate)
  void correction(const float* r, const float* invDiag, float* out = nullptr, bool accumulate = false);
  // pre-smoothing: x = alpha c(r); res = r - alpha A c when res != null (the
  // V-cycle residual comes free with the step)
  void smoothPre(deviceMemory<float>& o_r, deviceMemory<float>& o_x, deviceMemory<float>* o_resOut,
                 const float* invDiag, const op32_t& A, const res32_t& R);
  // one network step on a given residual: x += alpha c(res), res -= alpha A c
  void augment(deviceMemory<float>& o_x, deviceMemory<float>& o_res, const float* invDiag, const op32_t& A);
  // post-smoothing: res = r - A x, x += alpha c(res)
  void smoothPost(deviceMemory<float>& o_r, deviceMemory<float>& o_x, const float* invDiag, const op32_t& A,
                  const res32_t& R);

  // loss and gradient of one sample (r, target eh: FP32, gathered; *nrm = eh.A eh on
  // the device) into o_grad, then one Adam step when update is set
  void trainStep(const float* r, const float* eh, const double* nrm, const float* invDiag, const op32_t& A, float lr,
                 bool update = true);
  // run-time acceptance: in each solve the network stays in use while FPCG
  // keeps converging (finite residuals, no rejected step, and no two
  // consecutive iterations with |r| reduced by less than 10%); otherwise
  // Chebyshev takes over for the rest of that solve
  void monitorStart();
  void monitorIteration(int it, double rdotr);
  // ARCH_LINEAR least squares: add the Gram matrices of one sample (features
  // X of r, element targets T of eh) to G [nIn][nIn] += X^T X and
  // B [Np][nIn] += T^T X (FP32 GEMM on the device, FP64 sums on the host)
  void linearGramAccumulate(const float* r, const float* eh, const float* invDiag, std::vector<double>& G,
                            std::vector<double>& B);
  // W = B (G + ridge tr(G)/nIn I)^{-1} (FP64 Cholesky), omega = 0
  void linearSolve(const std::vector<double>& G, const std::vector<double>& B, double ridge);

  // statistics: [0] c.res, [1] c.Ac, [2] alpha, [3] rejected steps, [4] last loss, [5] loss sum
  std::vector<double> stats() const;
  void resetLossSum();
};

}  // namespace libp
