// Learned per-element smoother for the fine p-MG level (no reference counterpart).
#pragma once

#include "core.hpp"

namespace libp {

class elliptic_t;

/* A per-element MLP maps the element residual (Np values) and the six
 * face-neighbour residual traces (6 x Nfp) to an element error correction;
 * corrections are weighted by inverse multiplicity and assembled with the
 * gather-scatter. The weights are BF16 with FP32 masters; inference and
 * training run on the project's own MFMA kernels (libp_mlp_*). Used in place
 * of the Chebyshev pre/post smoothing on the fine level, with Chebyshev as
 * the fallback when no weights are loaded or the weights are rejected. */
class learnedSmoother_t {
 public:
  bool active = false;
  int Np = 0, Nfp = 0, nIn = 0, nHidden = 256, nOut = 0;
  int chebSteps = 0;  // Chebyshev sweeps applied after the network correction
  elliptic_t& elliptic;
  // layer l: W[l] (BF16, nIn_l x nOut_l row-major), b[l] (FP32)
  std::vector<deviceMemory<unsigned short>> W;
  std::vector<deviceMemory<float>> b;
  std::vector<deviceMemory<float>> master;  // FP32 masters of W (training)
  deviceMemory<int> o_traceIds;             // [E][6*Nfp] local id of the neighbour node, -1 at the boundary
  deviceMemory<float> o_feat, o_h1, o_h2, o_out, o_rL;

  explicit learnedSmoother_t(elliptic_t& el);
  void load(const std::string& path);
  void save(const std::string& path) const;
  void initRandom(unsigned seed);
  // x += S(r - A x) style smoothing step with the network (see learned.cpp)
  void smooth(elliptic_t& el, deviceMemory<float>& o_r, deviceMemory<float>& o_X, bool xIsZero,
              deviceMemory<float>& o_invDiag, deviceMemory<float>& o_s1, deviceMemory<float>& o_s2);
  // network features of a gathered FP32 residual: [E][nIn]
  void features(deviceMemory<float>& o_r);
  // forward pass on o_feat; output correction in o_out ([E][Np])
  void forward(int E);
};

}  // namespace libp
