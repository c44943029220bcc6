// Learned per-element smoother (no reference counterpart).
#include "learned.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "elliptic.hpp"

namespace libp {

namespace {

const char kMagic[8] = {'L', 'P', 'A', 'M', 'D', 'M', 'L', 'P'};
// version 1: MLP only; version 2 adds the architecture after the hidden width
const uint32_t kVersion = 2;

uint64_t fnv1a(const void* data, size_t n) {
  const unsigned char* p = (const unsigned char*)data;
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

// splitmix64 (Steele, Lea, Flood 2014): deterministic initial weights
struct splitmix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

learnedSmoother_t::learnedSmoother_t(elliptic_t& el, int _arch) : elliptic(el), platform(el.platform), arch(_arch) {
  const mesh_t& mesh = el.mesh;
  LIBP_ABORT("unknown learned smoother architecture " << arch, arch != ARCH_MLP && arch != ARCH_LINEAR);
  Np = mesh.Np;
  Nfp = mesh.Nfp;
  nIn = Np + 6 * Nfp;
  nOut = Np;
  E = mesh.Nelements;
  Ndofs = el.Ndofs;
  if (arch == ARCH_LINEAR) {
    H = 0;
    oW1 = ob1 = oW2 = ob2 = ob3 = 0;
    oW3 = 0;
    oOm = (size_t)nOut * nIn;
  } else {
    oW1 = 0;
    ob1 = oW1 + (size_t)H * nIn;
    oW2 = ob1 + H;
    ob2 = oW2 + (size_t)H * H;
    oW3 = ob2 + H;
    ob3 = oW3 + (size_t)nOut * H;
    oOm = ob3 + nOut;
  }
  nParams = oOm + 1;

  o_theta = platform.malloc<float>(nParams);
  o_grad = platform.malloc<float>(nParams);
  o_m = platform.malloc<float>(nParams);
  o_v = platform.malloc<float>(nParams);
  o_thetaBf = platform.malloc<unsigned short>(nParams);
  fusedLayer = arch == ARCH_LINEAR && Np == 512 && 6 * Nfp == 384;
  if (fusedLayer) o_Wpacked = platform.malloc<unsigned short>((size_t)nOut * nIn);
  o_X = platform.malloc<unsigned short>((size_t)E * nIn);
  if (H > 0) {
    o_H1 = platform.malloc<unsigned short>((size_t)E * H);
    o_H2 = platform.malloc<unsigned short>((size_t)E * H);
  }
  o_scale = platform.malloc<float>(E);
  o_Y = platform.malloc<float>((size_t)E * nOut);
  o_c = platform.malloc<float>(Ndofs);
  o_Ac = platform.malloc<float>(Ndofs);
  o_res = platform.malloc<float>(Ndofs);
  o_sc = platform.malloc<double>(8);
  o_rws = platform.malloc<char>(libp_reduce_ws_bytes());
  o_invLen = platform.malloc<float>(Ndofs);
  LIBP_HIP(hipMemsetAsync(o_sc.ptr(), 0, 8 * sizeof(double), platform.stream));
  LIBP_HIP(hipMemsetAsync(o_rws.ptr(), 0, libp_reduce_ws_bytes(), platform.stream));
  LIBP_CALL(libp_inv_row_length_f32(Ndofs, el.ogsMasked.o_rowStartsT.ptr(), o_invLen.ptr(), platform.s()));
  setupMaps();
  initRandom(1);
}

/* Neighbour traces from the box structure: element e = ex + nx (ey + ny ez),
 * node n = i + Nq (j + Nq k). Face f's nodes are taken in increasing n (the
 * reference's faceNodes order); the trace of face f is the neighbour's node
 * layer next to the shared face. Each face-node pair is checked to carry the
 * same global id. */
void learnedSmoother_t::setupMaps() {
  const mesh_t& mesh = elliptic.mesh;
  LIBP_ABORT("the learned smoother's neighbour traces need a BOX mesh", !mesh.boxMesh);
  const int Nq = mesh.Nq, N = mesh.N;
  const dlong nx = mesh.nx, ny = mesh.ny, nz = mesh.nz;
  std::vector<dlong> g2l((size_t)E * Np);
  elliptic.o_GlobalToLocal.copyTo(g2l.data(), g2l.size(), 0, platform.stream);
  platform.finish();
  std::vector<int> tr((size_t)E * 6 * Nfp, -1);
  auto node = [&](int i, int j, int k) { return i + Nq * (j + Nq * k); };
  for (dlong ez = 0; ez < nz; ++ez)
    for (dlong ey = 0; ey < ny; ++ey)
      for (dlong ex = 0; ex < nx; ++ex) {
        const dlong e = ex + nx * (ey + ny * ez);
        for (int f = 0; f < 6; ++f) {
          dlong qx = ex, qy = ey, qz = ez;
          if (f == 0) --qz;
          if (f == 1) --qy;
          if (f == 2) ++qx;
          if (f == 3) ++qy;
          if (f == 4) --qx;
          if (f == 5) ++qz;
          const bool out = qx < 0 || qy < 0 || qz < 0 || qx >= nx || qy >= ny || qz >= nz;
          if (out && !mesh.periodic) continue;
          qx = (qx + nx) % nx;
          qy = (qy + ny) % ny;
          qz = (qz + nz) % nz;
          const dlong en = qx + nx * (qy + ny * qz);
          for (int b = 0; b < Nq; ++b)
            for (int a = 0; a < Nq; ++a) {
              int me = 0, match = 0, inner = 0;
              switch (f) {
                case 0: me = node(a, b, 0); match = node(a, b, N); inner = node(a, b, N - 1); break;
                case 1: me = node(a, 0, b); match = node(a, N, b); inner = node(a, N - 1, b); break;
                case 2: me = node(N, a, b); match = node(0, a, b); inner = node(1, a, b); break;
                case 3: me = node(a, N, b); match = node(a, 0, b); inner = node(a, 1, b); break;
                case 4: me = node(0, a, b); match = node(N, a, b); inner = node(N - 1, a, b); break;
                default: me = node(a, b, N); match = node(a, b, 0); inner = node(a, b, 1); break;
              }
              LIBP_ABORT("learned smoother: face nodes of elements " << e << " and " << en << " do not match",
                         mesh.globalIds[(size_t)e * Np + me] != mesh.globalIds[(size_t)en * Np + match]);
              tr[((size_t)e * 6 + f) * Nfp + a + Nq * b] = g2l[(size_t)en * Np + inner];
            }
        }
      }
  o_trace = platform.malloc(tr);
}

void learnedSmoother_t::initRandom(unsigned long seed) {
  splitmix rng{seed};
  std::vector<float> th(nParams, 0.f);
  auto fill = [&](size_t o, int fanOut, int fanIn, double bound) {
    for (size_t i = 0; i < (size_t)fanOut * fanIn; ++i) th[o + i] = (float)((2 * rng.uniform() - 1) * bound);
  };
  if (arch == ARCH_MLP) {
    fill(oW1, H, nIn, std::sqrt(6.0 / nIn));
    fill(oW2, H, H, std::sqrt(6.0 / H));
  }
  // last layer starts at zero: the initial correction is omega D^{-1} r (damped Jacobi)
  th[oOm] = 0.5f;
  setParameters(th);
  active = false;
  status = "random initialisation";
}

std::vector<float> learnedSmoother_t::parameters() const {
  std::vector<float> th(nParams);
  o_theta.copyTo(th.data(), nParams, 0, platform.stream);
  platform.finish();
  return th;
}

void learnedSmoother_t::setParameters(const std::vector<float>& th) {
  LIBP_ABORT("learned smoother: parameter vector has " << th.size() << " entries, expected " << nParams,
             th.size() != nParams);
  o_theta.copyFrom(th.data(), nParams, 0, platform.stream);
  omegaHost = th[oOm];
  LIBP_CALL(libp_f32_to_bf16(nParams, o_theta.ptr(), o_thetaBf.ptr(), platform.s()));
  if (fusedLayer)  // the only writer of the linear weights (Adam trains the MLP only)
    LIBP_CALL(libp_learned_pack_weights_bf16(nOut, nIn, o_thetaBf.ptr() + oW3, nIn, o_Wpacked.ptr(), platform.s()));
  LIBP_HIP(hipMemsetAsync(o_m.ptr(), 0, nParams * sizeof(float), platform.stream));
  LIBP_HIP(hipMemsetAsync(o_v.ptr(), 0, nParams * sizeof(float), platform.stream));
  adamStep = 0;
  platform.finish();
}

bool learnedSmoother_t::load(const std::string& path) {
  active = false;
  fallback = false;
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    status = "rejected: cannot open " + path;
    return false;
  }
  char magic[8];
  uint32_t ver = 0, np = 0, nfp = 0, h = 0;
  uint64_t n = 0, hash = 0;
  f.read(magic, 8);
  f.read((char*)&ver, 4);
  f.read((char*)&np, 4);
  f.read((char*)&nfp, 4);
  f.read((char*)&h, 4);
  f.read((char*)&n, 8);
  f.read((char*)&hash, 8);
  uint32_t fileArch = ARCH_MLP;
  if (ver == 2) f.read((char*)&fileArch, 4);
  if (!f || std::memcmp(magic, kMagic, 8) != 0 || (ver != 1 && ver != 2)) {
    status = "rejected: not a learned-smoother weight file";
    return false;
  }
  if ((int)fileArch != arch) {
    status = "rejected: weights are for another architecture";
    return false;
  }
  if ((int)np != Np || (int)nfp != Nfp || (int)h != H || n != nParams) {
    status = "rejected: weights are for another shape";
    return false;
  }
  std::vector<float> th(nParams);
  f.read((char*)th.data(), nParams * sizeof(float));
  if (!f) {
    status = "rejected: truncated weight file";
    return false;
  }
  if (fnv1a(th.data(), nParams * sizeof(float)) != hash) {
    status = "rejected: hash mismatch";
    return false;
  }
  for (float v : th)
    if (!std::isfinite(v)) {
      status = "rejected: non-finite weights";
      return false;
    }
  setParameters(th);
  active = true;
  status = "loaded " + path;
  return true;
}

void learnedSmoother_t::save(const std::string& path) const {
  std::vector<float> th = parameters();
  std::ofstream f(path, std::ios::binary);
  LIBP_ABORT("learned smoother: cannot write " << path, !f);
  const uint32_t ver = kVersion, np = Np, nfp = Nfp, h = H, a = arch;
  const uint64_t n = nParams, hash = fnv1a(th.data(), nParams * sizeof(float));
  f.write(kMagic, 8);
  f.write((const char*)&ver, 4);
  f.write((const char*)&np, 4);
  f.write((const char*)&nfp, 4);
  f.write((const char*)&h, 4);
  f.write((const char*)&n, 8);
  f.write((const char*)&hash, 8);
  f.write((const char*)&a, 4);
  f.write((const char*)th.data(), nParams * sizeof(float));
  LIBP_ABORT("learned smoother: write failed for " << path, !f);
}

void learnedSmoother_t::correction(const float* r, const float* invDiag, float* out, bool accumulate) {
  void* s = platform.s();
  const unsigned short* Wb = o_thetaBf.ptr();
  const float* th = o_theta.ptr();
  // the output layer multiplies each element's row by its input scale s_e in its
  // epilogue, so the assembly below reads one value per copy (scale NULL)
  if (fusedLayer) {  // same Y to the bit as features + GEMM, without the input round trip
    LIBP_CALL(libp_learned_linear_layer_f32(E, Np, 6 * Nfp, nOut, elliptic.o_GlobalToLocal.ptr(), o_trace.ptr(), r,
                                            invDiag, o_Wpacked.ptr(), o_Y.ptr(), nOut, o_scale.ptr(), s));
  } else if (arch == ARCH_LINEAR) {
    LIBP_CALL(libp_learned_features_f32(E, Np, 6 * Nfp, elliptic.o_GlobalToLocal.ptr(), o_trace.ptr(), r, invDiag,
                                        o_X.ptr(), nIn, o_scale.ptr(), s));
    LIBP_CALL(libp_mlp_linear_fwd_scaled(E, nIn, nOut, o_X.ptr(), nIn, Wb + oW3, nIn, nullptr, o_scale.ptr(),
                                         o_Y.ptr(), nOut, s));
  } else {
    LIBP_CALL(libp_learned_features_f32(E, Np, 6 * Nfp, elliptic.o_GlobalToLocal.ptr(), o_trace.ptr(), r, invDiag,
                                        o_X.ptr(), nIn, o_scale.ptr(), s));
    LIBP_CALL(libp_mlp_linear_fwd(E, nIn, H, o_X.ptr(), nIn, Wb + oW1, nIn, th + ob1, 1, o_H1.ptr(), H, s));
    LIBP_CALL(libp_mlp_linear_fwd(E, H, H, o_H1.ptr(), H, Wb + oW2, H, th + ob2, 1, o_H2.ptr(), H, s));
    LIBP_CALL(libp_mlp_linear_fwd_scaled(E, H, nOut, o_H2.ptr(), H, Wb + oW3, H, th + ob3, o_scale.ptr(), o_Y.ptr(),
                                         nOut, s));
  }
  // the linear smoother's omega only changes through setParameters: skip the term when it is zero
  const float* om = (arch == ARCH_LINEAR && omegaHost == 0.f) ? nullptr : th + oOm;
  const int* rs = elliptic.ogsMasked.o_rowStartsT.ptr();
  const int* ci = elliptic.ogsMasked.o_colIdsT.ptr();
  if (accumulate)
    LIBP_CALL(libp_learned_assemble_add_f32(Ndofs, rs, ci, Np, o_Y.ptr(), nOut, nullptr, om, r, invDiag, out, s));
  else
    LIBP_CALL(libp_learned_assemble_f32(Ndofs, rs, ci, Np, o_Y.ptr(), nOut, nullptr, om, r, invDiag,
                                        out ? out : o_c.ptr(), s));
}

void learnedSmoother_t::smoothPre(deviceMemory<float>& o_r, deviceMemory<float>& o_x, deviceMemory<float>* o_resOut,
                                  const float* invDiag, const op32_t& A, const res32_t& R) {
  void* s = platform.s();
  if (arch == ARCH_LINEAR) {  // x = c(r), then x += c(r - A x) for further sweeps
    correction(o_r.ptr(), invDiag, o_x.ptr(), false);
    for (int k = 1; k < sweeps; ++k) {
      R(o_x, o_r, o_res);
      correction(o_res.ptr(), invDiag, o_x.ptr(), true);
    }
    if (o_resOut) R(o_x, o_r, *o_resOut);
    return;
  }
  correction(o_r.ptr(), invDiag);
  A(o_c, o_Ac);
  LIBP_CALL(libp_learned_alpha_f32(Ndofs, o_c.ptr(), o_r.ptr(), o_Ac.ptr(), o_sc.ptr(), o_rws.ptr(), s));
  float* res = nullptr;
  if (o_resOut) {
    LIBP_HIP(hipMemcpyAsync(o_resOut->ptr(), o_r.ptr(), Ndofs * sizeof(float), hipMemcpyDeviceToDevice,
                            platform.stream));
    res = o_resOut->ptr();
  }
  LIBP_CALL(libp_learned_update_f32(Ndofs, 1, o_sc.ptr(), o_c.ptr(), o_Ac.ptr(), o_x.ptr(), res, s));
}

void learnedSmoother_t::augment(deviceMemory<float>& o_x, deviceMemory<float>& o_resIO, const float* invDiag,
                                const op32_t& A) {
  void* s = platform.s();
  correction(o_resIO.ptr(), invDiag);
  A(o_c, o_Ac);
  if (arch == ARCH_LINEAR) {  // fixed step: x += c, res -= A c
    LIBP_CALL(libp_axpy_f32(Ndofs, 1.f, o_c.ptr(), 1.f, o_x.ptr(), s));
    LIBP_CALL(libp_axpy_f32(Ndofs, -1.f, o_Ac.ptr(), 1.f, o_resIO.ptr(), s));
    return;
  }
  LIBP_CALL(libp_learned_alpha_f32(Ndofs, o_c.ptr(), o_resIO.ptr(), o_Ac.ptr(), o_sc.ptr(), o_rws.ptr(), s));
  LIBP_CALL(libp_learned_update_f32(Ndofs, 0, o_sc.ptr(), o_c.ptr(), o_Ac.ptr(), o_x.ptr(), o_resIO.ptr(), s));
}

void learnedSmoother_t::smoothPost(deviceMemory<float>& o_r, deviceMemory<float>& o_x, const float* invDiag,
                                   const op32_t& A, const res32_t& R) {
  void* s = platform.s();
  if (arch == ARCH_LINEAR) {  // sweeps of x += c(r - A x)
    for (int k = 0; k < sweeps; ++k) {
      R(o_x, o_r, o_res);
      correction(o_res.ptr(), invDiag, o_x.ptr(), true);
    }
    return;
  }
  R(o_x, o_r, o_res);
  correction(o_res.ptr(), invDiag);
  A(o_c, o_Ac);
  LIBP_CALL(libp_learned_alpha_f32(Ndofs, o_c.ptr(), o_res.ptr(), o_Ac.ptr(), o_sc.ptr(), o_rws.ptr(), s));
  LIBP_CALL(libp_learned_update_f32(Ndofs, 0, o_sc.ptr(), o_c.ptr(), o_Ac.ptr(), o_x.ptr(), nullptr, s));
}

void learnedSmoother_t::trainStep(const float* r, const float* eh, const double* nrm, const float* invDiag,
                                  const op32_t& A, float lr, bool update) {
  LIBP_ABORT("the linear smoother is fitted by least squares (linearGramAccumulate, linearSolve)",
             arch == ARCH_LINEAR);
  void* s = platform.s();
  if (!o_dY.ptr()) {
    o_dY = platform.malloc<unsigned short>((size_t)E * nOut);
    o_dH1 = platform.malloc<unsigned short>((size_t)E * H);
    o_dH2 = platform.malloc<unsigned short>((size_t)E * H);
    o_gc = platform.malloc<float>(Ndofs);
    o_d = platform.malloc<float>(Ndofs);
    o_Ad = platform.malloc<float>(Ndofs);
    long ws = std::max(libp_mlp_ws_bytes(E, nIn, H), std::max(libp_mlp_ws_bytes(E, H, H), libp_mlp_ws_bytes(E, H, nOut)));
    o_ws = platform.malloc<char>(ws);
  }
  const unsigned short* Wb = o_thetaBf.ptr();
  float* g = o_grad.ptr();
  correction(r, invDiag);
  LIBP_HIP(hipMemcpyAsync(o_d.ptr(), eh, Ndofs * sizeof(float), hipMemcpyDeviceToDevice, platform.stream));
  LIBP_CALL(libp_axpy_f32(Ndofs, -1.f, o_c.ptr(), 1.f, o_d.ptr(), s));
  A(o_d, o_Ad);
  LIBP_CALL(libp_learned_loss_f32(Ndofs, o_d.ptr(), o_Ad.ptr(), r, invDiag, nrm, o_gc.ptr(), o_sc.ptr() + 4, g + oOm,
                                  o_rws.ptr(), s));
  LIBP_CALL(libp_learned_scatter_grad_f32(E, Np, elliptic.o_GlobalToLocal.ptr(), o_gc.ptr(), o_invLen.ptr(),
                                          o_scale.ptr(), o_dY.ptr(), nOut, s));
  void* ws = o_ws.ptr();
  LIBP_CALL(libp_mlp_linear_bwd_weight(E, H, nOut, o_dY.ptr(), nOut, o_H2.ptr(), H, g + oW3, g + ob3, ws, s));
  LIBP_CALL(libp_mlp_linear_bwd_data(E, H, nOut, o_dY.ptr(), nOut, Wb + oW3, H, o_H2.ptr(), H, o_dH2.ptr(), H, s));
  LIBP_CALL(libp_mlp_linear_bwd_weight(E, H, H, o_dH2.ptr(), H, o_H1.ptr(), H, g + oW2, g + ob2, ws, s));
  LIBP_CALL(libp_mlp_linear_bwd_data(E, H, H, o_dH2.ptr(), H, Wb + oW2, H, o_H1.ptr(), H, o_dH1.ptr(), H, s));
  LIBP_CALL(libp_mlp_linear_bwd_weight(E, nIn, H, o_dH1.ptr(), H, o_X.ptr(), nIn, g + oW1, g + ob1, ws, s));
  if (update)
    LIBP_CALL(libp_adam_step((int)nParams, o_theta.ptr(), o_thetaBf.ptr(), g, o_m.ptr(), o_v.ptr(), lr, 0.9f, 0.999f,
                             1e-8f, ++adamStep, s));
}

void learnedSmoother_t::linearGramAccumulate(const float* r, const float* eh, const float* invDiag,
                                             std::vector<double>& G, std::vector<double>& B) {
  LIBP_ABORT("least squares fits the linear smoother only", arch != ARCH_LINEAR);
  void* s = platform.s();
  if (!o_T.ptr()) {
    o_T = platform.malloc<unsigned short>((size_t)E * nOut);
    o_grad = platform.malloc<float>(std::max<size_t>(nParams, (size_t)nIn * nIn));
    o_d = platform.malloc<float>(nIn);
    o_ws = platform.malloc<char>(std::max(libp_mlp_ws_bytes(E, nIn, nIn), libp_mlp_ws_bytes(E, nIn, nOut)));
  }
  G.resize((size_t)nIn * nIn, 0.0);
  B.resize((size_t)nOut * nIn, 0.0);
  LIBP_CALL(libp_learned_features_f32(E, Np, 6 * Nfp, elliptic.o_GlobalToLocal.ptr(), o_trace.ptr(), r, invDiag,
                                      o_X.ptr(), nIn, o_scale.ptr(), s));
  LIBP_CALL(libp_learned_element_target_f32(E, Np, elliptic.o_GlobalToLocal.ptr(), eh, o_scale.ptr(), o_T.ptr(), nOut,
                                            s));
  std::vector<float> h;
  auto addTo = [&](std::vector<double>& acc, size_t n) {
    h.resize(n);
    o_grad.copyTo(h.data(), n, 0, platform.stream);
    platform.finish();
    for (size_t i = 0; i < n; ++i) acc[i] += h[i];
  };
  // G[j][k] = sum_e X[e][j] X[e][k] ; B[n][k] = sum_e T[e][n] X[e][k]
  LIBP_CALL(libp_mlp_linear_bwd_weight(E, nIn, nIn, o_X.ptr(), nIn, o_X.ptr(), nIn, o_grad.ptr(), o_d.ptr(),
                                       o_ws.ptr(), s));
  addTo(G, (size_t)nIn * nIn);
  LIBP_CALL(libp_mlp_linear_bwd_weight(E, nIn, nOut, o_T.ptr(), nOut, o_X.ptr(), nIn, o_grad.ptr(), o_d.ptr(),
                                       o_ws.ptr(), s));
  addTo(B, (size_t)nOut * nIn);
}

void learnedSmoother_t::linearSolve(const std::vector<double>& G, const std::vector<double>& B, double ridge) {
  LIBP_ABORT("least squares fits the linear smoother only", arch != ARCH_LINEAR);
  const int n = nIn;
  LIBP_ABORT("Gram matrices of the wrong size", G.size() != (size_t)n * n || B.size() != (size_t)nOut * n);
  double tr = 0;
  for (int i = 0; i < n; ++i) tr += G[(size_t)i * n + i];
  LIBP_ABORT("empty Gram matrix (no samples)", !(tr > 0));
  // Cholesky L L^T of the ridge-regularised (symmetrised) Gram matrix
  std::vector<double> L((size_t)n * n, 0.0);
  const double shift = ridge * tr / n;
  for (int j = 0; j < n; ++j) {
    double d = 0.5 * (G[(size_t)j * n + j] + G[(size_t)j * n + j]) + shift;
    for (int k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    LIBP_ABORT("Gram matrix not positive definite", !(d > 0));
    const double ljj = std::sqrt(d);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double v = 0.5 * (G[(size_t)i * n + j] + G[(size_t)j * n + i]);
      for (int k = 0; k < j; ++k) v -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = v / ljj;
    }
  }
  // row o of W solves (G + shift I) w = B[o]
  std::vector<float> th(nParams, 0.f);
  std::vector<double> y(n);
  for (int o = 0; o < nOut; ++o) {
    for (int i = 0; i < n; ++i) {
      double v = B[(size_t)o * n + i];
      for (int k = 0; k < i; ++k) v -= L[(size_t)i * n + k] * y[k];
      y[i] = v / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double v = y[i];
      for (int k = i + 1; k < n; ++k) v -= L[(size_t)k * n + i] * y[k];
      y[i] = v / L[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) th[oW3 + (size_t)o * n + i] = (float)y[i];
  }
  th[oOm] = 0.f;
  setParameters(th);
}

void learnedSmoother_t::monitorStart() {
  monitorHist.clear();
  if (!active) return;
  fallback = false;  // every solve starts with the network
  // only the MLP's energy-optimal step can be rejected: the linear smoother's
  // fixed step needs no device read (and no stream synchronisation) here
  monitorRejected = (arch == ARCH_MLP) ? stats()[3] : 0.0;
}

void learnedSmoother_t::monitorIteration(int it, double rdotr) {
  (void)it;
  if (!active || fallback) return;
  const double rej = (arch == ARCH_MLP) ? stats()[3] : 0.0;
  std::string why;
  if (!std::isfinite(rdotr)) why = "non-finite residual";
  if (rej > monitorRejected) why = "rejected step";
  monitorHist.push_back(rdotr);
  const size_t n = monitorHist.size();
  if (n >= 3 && monitorHist[n - 1] > 0.81 * monitorHist[n - 2] && monitorHist[n - 2] > 0.81 * monitorHist[n - 3])
    why = "residual reduction below 10% twice";
  if (!why.empty()) {
    fallback = true;
    rejectedSteps = (int)rej;
    status = "fallback to Chebyshev: " + why;
  }
}

std::vector<double> learnedSmoother_t::stats() const {
  std::vector<double> h(8);
  o_sc.copyTo(h.data(), 8, 0, platform.stream);
  platform.finish();
  return h;
}

void learnedSmoother_t::resetLossSum() {
  const double z = 0;
  o_sc.copyFrom(&z, 1, 5, platform.stream);
  platform.finish();
}

}  // namespace libp
