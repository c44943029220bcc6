// Learned per-element smoother (no reference counterpart).
#include "learned.hpp"

#include "elliptic.hpp"

namespace libp {

learnedSmoother_t::learnedSmoother_t(elliptic_t& el) : elliptic(el) {
  Np = el.mesh.Np;
  Nfp = el.mesh.Nfp;
}

void learnedSmoother_t::load(const std::string&) { LIBP_FORCE_ABORT("learned smoother weights: not available"); }
void learnedSmoother_t::save(const std::string&) const {}
void learnedSmoother_t::initRandom(unsigned) {}
void learnedSmoother_t::features(deviceMemory<float>&) {}
void learnedSmoother_t::forward(int) {}
void learnedSmoother_t::smooth(elliptic_t&, deviceMemory<float>&, deviceMemory<float>&, bool, deviceMemory<float>&,
                               deviceMemory<float>&, deviceMemory<float>&) {
  LIBP_FORCE_ABORT("learned smoother not active");
}

}  // namespace libp
