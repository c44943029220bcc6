/* CPU oracle of the learned per-element smoother (test infrastructure only).
 *
 * The learned smoother has no counterpart in the reference, so this file
 * restates the build's own definition of it, independently of the GPU code:
 *   - neighbour traces from the reference's face connectivity (EToE, EToF,
 *     vmapP, faceNodes of meshConnect*.cpp / meshConnectFaceNodes3D.cpp as
 *     restated in mesh.cpp), where the GPU code uses the box structure;
 *   - the network input, forward pass, assembly, energy loss and the full
 *     backward pass in double precision, optionally rounding to BF16 at the
 *     points where the GPU kernels store BF16 (so the GPU can be compared
 *     closely), and without rounding for finite-difference gradient checks.
 *
 * Parameter vector theta (FP32), for nIn = Np + 6 Nfp, H hidden, nOut = Np:
 *   W1 [H][nIn], b1 [H], W2 [H][H], b2 [H], W3 [nOut][H], b3 [nOut], omega.
 * H = 0 is the linear architecture: W [nOut][nIn] (no bias), omega; its
 * least-squares fit is restated by LearnedLinearGram.
 */
#include <cmath>
#include <cstring>

#include "oracle.hpp"

namespace oracle {

struct LearnedDims {
  int nIn, H, nOut;
  size_t oW1, ob1, oW2, ob2, oW3, ob3, oOm, n;
};

LearnedDims learnedDims(int Np, int Nfp, int H) {
  LearnedDims d;
  d.nIn = Np + 6 * Nfp;
  d.H = H;
  d.nOut = Np;
  if (H == 0) {
    d.oW1 = d.ob1 = d.oW2 = d.ob2 = d.oW3 = d.ob3 = 0;
    d.oOm = (size_t)d.nOut * d.nIn;
    d.n = d.oOm + 1;
    return d;
  }
  d.oW1 = 0;
  d.ob1 = d.oW1 + (size_t)H * d.nIn;
  d.oW2 = d.ob1 + H;
  d.ob2 = d.oW2 + (size_t)H * H;
  d.oW3 = d.ob2 + H;
  d.ob3 = d.oW3 + (size_t)d.nOut * H;
  d.oOm = d.ob3 + d.nOut;
  d.n = d.oOm + 1;
  return d;
}

// round to the nearest BF16 value (ties to even)
double bf16r(double x) {
  float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return x;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

/* traceIds[e][f*Nfp + m]: gathered id of the neighbour node one layer inside
 * the neighbour across face f, matched to this element's m-th face node
 * (faceNodes order); -1 without a neighbour or on a Dirichlet node. */
void LearnedTraceIds(const Elliptic& el, std::vector<int>& trace) {
  const Mesh& m = el.mesh;
  const int Nq = m.Nq, Nfp = m.Nfp;
  // inward step (in local node index) from a node on face f of an element
  const int step[6] = {Nq * Nq, Nq, -1, -Nq, 1, -Nq * Nq};
  trace.assign((size_t)m.Nelements * 6 * Nfp, -1);
  for (dlong e = 0; e < m.Nelements; ++e)
    for (int f = 0; f < 6; ++f) {
      const dlong eP = m.EToE[e * 6 + f];
      if (eP < 0) continue;
      const int fP = m.EToF[e * 6 + f];
      for (int n = 0; n < Nfp; ++n) {
        const dlong idP = m.vmapP[((size_t)e * 6 + f) * Nfp + n];  // volume id in eP on the shared face
        const dlong inner = idP + step[fP];
        trace[((size_t)e * 6 + f) * Nfp + n] = (int)el.GlobalToLocal[inner];
      }
    }
}

struct Fwd {
  std::vector<double> X, scale, H1, H2, Y, c;
};

static void forward(const Elliptic& el, const std::vector<int>& trace, const LearnedDims& d, const double* th,
                    const double* r, const double* invD, bool bf, Fwd& F) {
  const Mesh& m = el.mesh;
  const dlong E = m.Nelements;
  const int Np = m.Np, Ntr = 6 * m.Nfp, nIn = d.nIn, H = d.H, nOut = d.nOut;
  auto R = [&](double x) { return bf ? bf16r(x) : x; };
  F.X.assign((size_t)E * nIn, 0.0);
  F.scale.assign(E, 0.0);
  for (dlong e = 0; e < E; ++e) {
    double ss = 0;
    std::vector<double> z(nIn);
    for (int f = 0; f < nIn; ++f) {
      const long g = (f < Np) ? (long)el.GlobalToLocal[(size_t)e * Np + f] : (long)trace[(size_t)e * Ntr + f - Np];
      // the GPU forms z in FP32 from FP32 r and invDiag
      z[f] = (g >= 0) ? (bf ? (double)((float)invD[g] * (float)r[g]) : invD[g] * r[g]) : 0.0;
      ss += z[f] * z[f];
    }
    const double s = std::sqrt(ss / nIn);
    F.scale[e] = bf ? (double)(float)s : s;
    const double inv = s > 0 ? 1.0 / F.scale[e] : 0.0;
    for (int f = 0; f < nIn; ++f) F.X[(size_t)e * nIn + f] = R(z[f] * inv);
  }
  auto W = [&](size_t o) { return R(th[o]); };
  if (H == 0) {  // linear: y_e = W x_e
    F.Y.assign((size_t)E * nOut, 0.0);
    std::vector<double> Wl((size_t)nOut * nIn);
    for (size_t i = 0; i < Wl.size(); ++i) Wl[i] = W(d.oW3 + i);
#pragma omp parallel for schedule(static)
    for (dlong e = 0; e < E; ++e) {
      const double* x = &F.X[(size_t)e * nIn];
      for (int o = 0; o < nOut; ++o) {
        double a = 0;
        for (int f = 0; f < nIn; ++f) a += Wl[(size_t)o * nIn + f] * x[f];
        F.Y[(size_t)e * nOut + o] = bf ? (double)(float)a : a;
      }
    }
  }
  if (H > 0) {
    F.H1.assign((size_t)E * H, 0.0);
    F.H2.assign((size_t)E * H, 0.0);
    F.Y.assign((size_t)E * nOut, 0.0);
    std::vector<double> W1((size_t)H * nIn), W2((size_t)H * H), W3((size_t)nOut * H);
    for (size_t i = 0; i < W1.size(); ++i) W1[i] = W(d.oW1 + i);
    for (size_t i = 0; i < W2.size(); ++i) W2[i] = W(d.oW2 + i);
    for (size_t i = 0; i < W3.size(); ++i) W3[i] = W(d.oW3 + i);
#pragma omp parallel for schedule(static)
    for (dlong e = 0; e < E; ++e) {
      const double* x = &F.X[(size_t)e * nIn];
      for (int h = 0; h < H; ++h) {
        double a = th[d.ob1 + h];
        for (int f = 0; f < nIn; ++f) a += W1[(size_t)h * nIn + f] * x[f];
        F.H1[(size_t)e * H + h] = R(a > 0 ? a : 0.0);
      }
      for (int h = 0; h < H; ++h) {
        double a = th[d.ob2 + h];
        for (int k = 0; k < H; ++k) a += W2[(size_t)h * H + k] * F.H1[(size_t)e * H + k];
        F.H2[(size_t)e * H + h] = R(a > 0 ? a : 0.0);
      }
      for (int o = 0; o < nOut; ++o) {
        double a = th[d.ob3 + o];
        for (int k = 0; k < H; ++k) a += W3[(size_t)o * H + k] * F.H2[(size_t)e * H + k];
        F.Y[(size_t)e * nOut + o] = bf ? (double)(float)a : a;
      }
    }
  }
  // assembly over the rows of the masked gather
  const OgsOperator& G = el.ogsMasked.gatherLocal;
  F.c.assign(el.Ndofs, 0.0);
  const double om = th[d.oOm];
  for (dlong g = 0; g < el.Ndofs; ++g) {
    const dlong a = G.rowStartsT[g], b = G.rowStartsT[g + 1];
    double acc = 0;
    for (dlong j = a; j < b; ++j) {
      const dlong id = G.colIdsT[j];
      const dlong e = id / Np;
      acc += F.scale[e] * F.Y[(size_t)e * nOut + (id - e * Np)];
    }
    F.c[g] = (b > a ? acc / (double)(b - a) : 0.0) + om * invD[g] * r[g];
  }
}

/* Loss J = (eh - c).A(eh - c) / nrm of one sample and dJ/dtheta (grad, length
 * learnedDims(...).n). Returns J; c (length Ndofs) if non-null. */
double LearnedLossGrad(const Elliptic& el, const std::vector<int>& trace, int H, const double* th, const double* r,
                       const double* eh, double nrm, const double* invD, bool bf, double* grad, double* cOut) {
  const Mesh& m = el.mesh;
  const dlong E = m.Nelements;
  const int Np = m.Np;
  const LearnedDims d = learnedDims(Np, m.Nfp, H);
  auto R = [&](double x) { return bf ? bf16r(x) : x; };
  Fwd F;
  forward(el, trace, d, th, r, invD, bf, F);
  if (cOut) std::memcpy(cOut, F.c.data(), sizeof(double) * el.Ndofs);
  std::vector<double> dv(el.Ndofs), Ad(el.Ndofs), scratch;
  for (dlong g = 0; g < el.Ndofs; ++g) dv[g] = eh[g] - F.c[g];
  el.Operator(dv.data(), Ad.data(), scratch);
  double J = 0;
  for (dlong g = 0; g < el.Ndofs; ++g) J += dv[g] * Ad[g];
  J /= nrm;
  if (!grad) return J;
  std::memset(grad, 0, sizeof(double) * d.n);
  std::vector<double> gc(el.Ndofs);
  double gOm = 0;
  for (dlong g = 0; g < el.Ndofs; ++g) {
    gc[g] = -2.0 * Ad[g] / nrm;
    gOm += gc[g] * invD[g] * r[g];
  }
  grad[d.oOm] = gOm;
  // dJ/dY[e][n] = gc[g] scale_e / |row g|
  const OgsOperator& G = el.ogsMasked.gatherLocal;
  std::vector<double> invLen(el.Ndofs);
  for (dlong g = 0; g < el.Ndofs; ++g) {
    const dlong len = G.rowStartsT[g + 1] - G.rowStartsT[g];
    invLen[g] = len > 0 ? 1.0 / len : 0.0;
  }
  const int nIn = d.nIn, nOut = d.nOut;
  std::vector<double> dY((size_t)E * nOut, 0.0), dH2((size_t)E * H), dH1((size_t)E * H);
  for (dlong e = 0; e < E; ++e)
    for (int n = 0; n < Np; ++n) {
      const long g = el.GlobalToLocal[(size_t)e * Np + n];
      dY[(size_t)e * nOut + n] = (g >= 0) ? R(gc[g] * F.scale[e] * invLen[g]) : 0.0;
    }
  auto W = [&](size_t o) { return R(th[o]); };
#pragma omp parallel for schedule(static)
  for (dlong e = 0; e < E; ++e) {
    for (int k = 0; k < H; ++k) {
      double a = 0;
      for (int o = 0; o < nOut; ++o) a += dY[(size_t)e * nOut + o] * W(d.oW3 + (size_t)o * H + k);
      dH2[(size_t)e * H + k] = (F.H2[(size_t)e * H + k] > 0) ? R(a) : 0.0;
    }
    for (int k = 0; k < H; ++k) {
      double a = 0;
      for (int h = 0; h < H; ++h) a += dH2[(size_t)e * H + h] * W(d.oW2 + (size_t)h * H + k);
      dH1[(size_t)e * H + k] = (F.H1[(size_t)e * H + k] > 0) ? R(a) : 0.0;
    }
  }
  // weight and bias gradients: sums over elements
#pragma omp parallel for schedule(static)
  for (int o = 0; o < nOut; ++o) {
    double gb = 0;
    std::vector<double> gw(H, 0.0);
    for (dlong e = 0; e < E; ++e) {
      const double v = dY[(size_t)e * nOut + o];
      gb += v;
      if (v != 0)
        for (int k = 0; k < H; ++k) gw[k] += v * F.H2[(size_t)e * H + k];
    }
    grad[d.ob3 + o] = gb;
    for (int k = 0; k < H; ++k) grad[d.oW3 + (size_t)o * H + k] = gw[k];
  }
#pragma omp parallel for schedule(static)
  for (int h = 0; h < H; ++h) {
    double gb2 = 0, gb1 = 0;
    std::vector<double> gw2(H, 0.0), gw1(nIn, 0.0);
    for (dlong e = 0; e < E; ++e) {
      const double v2 = dH2[(size_t)e * H + h], v1 = dH1[(size_t)e * H + h];
      gb2 += v2;
      gb1 += v1;
      if (v2 != 0)
        for (int k = 0; k < H; ++k) gw2[k] += v2 * F.H1[(size_t)e * H + k];
      if (v1 != 0)
        for (int f = 0; f < nIn; ++f) gw1[f] += v1 * F.X[(size_t)e * nIn + f];
    }
    grad[d.ob2 + h] = gb2;
    grad[d.ob1 + h] = gb1;
    for (int k = 0; k < H; ++k) grad[d.oW2 + (size_t)h * H + k] = gw2[k];
    for (int f = 0; f < nIn; ++f) grad[d.oW1 + (size_t)h * nIn + f] = gw1[f];
  }
  return J;
}

/* Gram matrices of the linear smoother's least-squares fit for one sample:
 * X = network input of r (BF16-rounded when bf), T[e][n] = eh[g] / s_e at the
 * element's nodes (0 at masked nodes, BF16-rounded when bf);
 * G [nIn][nIn] = X^T X, B [Np][nIn] = T^T X (sums over elements). */
void LearnedLinearGram(const Elliptic& el, const std::vector<int>& trace, const double* r, const double* eh,
                       const double* invD, bool bf, double* Gout, double* Bout) {
  const Mesh& m = el.mesh;
  const dlong E = m.Nelements;
  const int Np = m.Np;
  const LearnedDims d = learnedDims(Np, m.Nfp, 0);
  std::vector<double> th(d.n, 0.0);
  Fwd F;
  forward(el, trace, d, th.data(), r, invD, bf, F);
  const int nIn = d.nIn;
  std::vector<double> T((size_t)E * Np, 0.0);
  for (dlong e = 0; e < E; ++e)
    for (int n = 0; n < Np; ++n) {
      const long g = el.GlobalToLocal[(size_t)e * Np + n];
      const double s = F.scale[e];
      if (g < 0 || !(s > 0)) continue;
      // the GPU divides in FP32 (s is already FP32 there) and stores BF16
      T[(size_t)e * Np + n] = bf ? bf16r((double)((float)eh[g] / (float)s)) : eh[g] / s;
    }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < nIn; ++j) {
    for (int k = 0; k < nIn; ++k) {
      double a = 0;
      for (dlong e = 0; e < E; ++e) a += F.X[(size_t)e * nIn + j] * F.X[(size_t)e * nIn + k];
      Gout[(size_t)j * nIn + k] = a;
    }
  }
#pragma omp parallel for schedule(static)
  for (int o = 0; o < Np; ++o) {
    for (int k = 0; k < nIn; ++k) {
      double a = 0;
      for (dlong e = 0; e < E; ++e) a += T[(size_t)e * Np + o] * F.X[(size_t)e * nIn + k];
      Bout[(size_t)o * nIn + k] = a;
    }
  }
}

/* Network input and correction only (no loss): X [E][nIn], scale [E], Y [E][Np], c [Ndofs]. */
void LearnedForward(const Elliptic& el, const std::vector<int>& trace, int H, const double* th, const double* r,
                    const double* invD, bool bf, double* X, double* scale, double* Y, double* c) {
  const LearnedDims d = learnedDims(el.mesh.Np, el.mesh.Nfp, H);
  Fwd F;
  forward(el, trace, d, th, r, invD, bf, F);
  if (X) std::memcpy(X, F.X.data(), sizeof(double) * F.X.size());
  if (scale) std::memcpy(scale, F.scale.data(), sizeof(double) * F.scale.size());
  if (Y) std::memcpy(Y, F.Y.data(), sizeof(double) * F.Y.size());
  if (c) std::memcpy(c, F.c.data(), sizeof(double) * F.c.size());
}

}  // namespace oracle
