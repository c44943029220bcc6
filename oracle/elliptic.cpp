// CPU oracle (test infrastructure only): C0 hex elliptic operator, boundary
// setup, diagonal, right-hand side, restated from solvers/elliptic/.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

static const double PI_ = 3.14159265358979323846;

/* Forcing and boundary data. "Sine3D" restates data/ellipticSine3D.h:27-51;
 * "Fourier3D" is a seeded 64-mode right-hand side with zero Dirichlet data
 * (no reference counterpart). */
struct DataFn {
  int kind = 0;  // 0 sine, 1 fourier
  double amp[64], ph[64];
  int kx[64], ky[64], kz[64];
  void setup(const Settings& s) {
    std::string f = s.get("DATA FILE");
    if (f.find("Fourier3D") != std::string::npos) {
      kind = 1;
      long seed = s.has("DATA SEED") ? s.geti("DATA SEED") : 1;
      unsigned short xs[3] = {0x330E, (unsigned short)(seed & 0xffff), (unsigned short)((seed >> 16) & 0xffff)};
      for (int m = 0; m < 64; ++m) {
        kx[m] = 1 + (int)(4 * erand48(xs));
        ky[m] = 1 + (int)(4 * erand48(xs));
        kz[m] = 1 + (int)(4 * erand48(xs));
        double k2 = kx[m] * kx[m] + ky[m] * ky[m] + kz[m] * kz[m];
        double sg = (erand48(xs) < 0.5) ? -1.0 : 1.0;
        amp[m] = sg / k2;
        ph[m] = 2 * PI_ * erand48(xs);
      }
    } else if (f.find("Sine3D") != std::string::npos) {
      kind = 0;
    } else {
      throw std::runtime_error("unsupported DATA FILE " + f);
    }
  }
  double forcing(double x, double y, double z, double lambda) const {
    if (kind == 0) return (3 * PI_ * PI_ + lambda) * std::sin(PI_ * x) * std::sin(PI_ * y) * std::sin(PI_ * z);
    double f = 0;
    for (int m = 0; m < 64; ++m) f += amp[m] * std::sin(PI_ * (kx[m] * x + ky[m] * y + kz[m] * z) + ph[m]);
    return f;
  }
  double dirichlet(double x, double y, double z) const {
    if (kind == 0) return std::sin(PI_ * x) * std::sin(PI_ * y) * std::sin(PI_ * z);
    return 0.0;
  }
};

/* elliptic_t::Setup (ellipticSetup.cpp:30-180), C0 branch. */
void Elliptic::Setup(const Mesh& m, const Settings& s, double lam, GlibcRand& rng) {
  mesh = m;
  settings = s;
  lambda = lam;
  BoundarySetup(rng);
  Ndofs = ogsMasked.Ngather;
}

/* ellipticBoundarySetup.cpp:30-140 with BCType = {0,1,2}
 * (ellipticMain.cpp:62-67). */
void Elliptic::BoundarySetup(GlibcRand& rng) {
  const int BCType[3] = {0, 1, 2};
  allNeumann = (lambda == 0) ? 1 : 0;
  allNeumannPenalty = 1.;
  EToB.assign((size_t)mesh.Nelements * mesh.Nfaces, 0);
  for (dlong e = 0; e < mesh.Nelements; e++)
    for (int f = 0; f < mesh.Nfaces; f++) {
      int bc = mesh.EToB[e * mesh.Nfaces + f];
      if (bc > 0) {
        int BC = BCType[bc];
        EToB[e * mesh.Nfaces + f] = BC;
        if (BC != 2) allNeumann = 0;
      }
    }
  const dlong NL = mesh.Nelements * mesh.Np;
  Nmasked = 0;
  mapB.assign(NL, 0);
  for (dlong n = 0; n < NL; n++) {
    int bc = mesh.mapB[n];
    if (bc > 0) {
      mapB[n] = BCType[bc];
      if (mapB[n] == 1) Nmasked++;
    }
  }
  maskIds.resize(Nmasked);
  Nmasked = 0;
  for (dlong n = 0; n < NL; n++)
    if (mapB[n] == 1) maskIds[Nmasked++] = n;
  maskedGlobalIds = mesh.globalIds;
  for (dlong n = 0; n < Nmasked; n++) maskedGlobalIds[maskIds[n]] = 0;
  ogsMasked.Setup(NL, maskedGlobalIds, true, rng);
  allNeumannScale = 1. / std::sqrt((double)ogsMasked.Ngather);
  weight.assign(NL, 1.0);
  weightG.assign(ogsMasked.Ngather, 0.0);
  ogsMasked.gatherLocal.GatherAdd(weightG.data(), weight.data(), true);
  for (dlong n = 0; n < ogsMasked.Ngather; ++n)
    if (weightG[n] > 0.0) weightG[n] = 1. / weightG[n];
  ogsMasked.gatherLocal.Scatter(weight.data(), weightG.data(), false);
  ogsMasked.SetupGlobalToLocalMapping(GlobalToLocal);
}

/* ellipticPartialAxHex3D_v0 (okl/ellipticAxHex3D.okl:156-295) in OCCA Serial
 * form: each @inner block runs to completion before the next. Covers every
 * element (one rank: localGatherElementList is 0..E-1, split in two halves by
 * ellipticOperator.cpp:42-104; the split does not change any value).
 * parallel: elements are shared among OpenMP threads, as the reference's
 * OpenMP mode runs the @outer loop; every element's values are the same. */
void Elliptic::PartialAx(const double* q, double* Aq, bool parallel) const {
  const int Nq = mesh.Nq, Np = mesh.Np, NqNq = Nq * Nq;
  const double* DT = mesh.D.data();
#pragma omp parallel if (parallel)
  {
  std::vector<double> s_q(NqNq), s_Gqr(NqNq), s_Gqs(NqNq);
  std::vector<double> r_q(NqNq * Nq), r_Aq(NqNq * Nq);
  std::vector<double> r_qt(NqNq), r_Gqt(NqNq), r_Auk(NqNq), G(7 * NqNq);
#pragma omp for schedule(static)
  for (dlong e = 0; e < mesh.Nelements; ++e) {
    const dlong element = mesh.localGatherElementList[e];
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        const size_t base = i + j * Nq + (size_t)element * Np;
        for (int k = 0; k < Nq; k++) {
          const dlong id = GlobalToLocal[base + k * NqNq];
          r_q[(j * Nq + i) * Nq + k] = (id != -1) ? q[id] : 0.0;
          r_Aq[(j * Nq + i) * Nq + k] = 0.0;
        }
      }
    for (int k = 0; k < Nq; k++) {
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const size_t gbase = (size_t)element * 6 * Np + k * NqNq + j * Nq + i;
          double* g = &G[7 * (j * Nq + i)];
          for (int c = 0; c < 6; ++c) g[c] = mesh.ggeo[gbase + c * Np];
          g[6] = mesh.wJ[(size_t)element * Np + k * NqNq + j * Nq + i];
        }
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const double* rq = &r_q[(j * Nq + i) * Nq];
          s_q[j * Nq + i] = rq[k];
          double qt = 0;
          for (int m = 0; m < Nq; m++) qt += DT[k * Nq + m] * rq[m];
          r_qt[j * Nq + i] = qt;
        }
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          double qr = 0, qs = 0;
          for (int m = 0; m < Nq; m++) {
            qr += DT[i * Nq + m] * s_q[j * Nq + m];
            qs += DT[j * Nq + m] * s_q[m * Nq + i];
          }
          const double* g = &G[7 * (j * Nq + i)];
          const double qt = r_qt[j * Nq + i];
          s_Gqs[j * Nq + i] = (g[1] * qr + g[3] * qs + g[4] * qt);
          s_Gqr[j * Nq + i] = (g[0] * qr + g[1] * qs + g[2] * qt);
          r_Gqt[j * Nq + i] = (g[2] * qr + g[4] * qs + g[5] * qt);
          r_Auk[j * Nq + i] = g[6] * lambda * r_q[(j * Nq + i) * Nq + k];
        }
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          double Auk = r_Auk[j * Nq + i];
          double* rA = &r_Aq[(j * Nq + i) * Nq];
          const double Gqt = r_Gqt[j * Nq + i];
          for (int m = 0; m < Nq; m++) {
            Auk += DT[m * Nq + j] * s_Gqs[m * Nq + i];
            rA[m] += DT[k * Nq + m] * Gqt;
            Auk += DT[m * Nq + i] * s_Gqr[j * Nq + m];
          }
          rA[k] += Auk;
        }
    }
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i)
        for (int k = 0; k < Nq; k++)
          Aq[(size_t)element * Np + k * NqNq + j * Nq + i] = r_Aq[(j * Nq + i) * Nq + k];
  }
  }
}

/* elliptic_t::Operator, C0 (ellipticOperator.cpp:29-107): Aq = Q^T A_L Q q. */
void Elliptic::Operator(const double* q, double* Aq, std::vector<double>& AqL) const {
  AqL.resize((size_t)mesh.Nelements * mesh.Np);
  PartialAx(q, AqL.data(), parallel);
  ogsMasked.gatherLocal.GatherAdd(Aq, AqL.data(), true, parallel);
}

/* ellipticBuildOperatorDiagonal.cpp:53-78 + :998-1057 (hex C0). */
void Elliptic::BuildOperatorDiagonal(std::vector<double>& diagA) const {
  const int Nq = mesh.Nq, Np = mesh.Np;
  const double* D = mesh.D.data();
  std::vector<double> A((size_t)mesh.Nelements * Np);
  for (dlong eM = 0; eM < mesh.Nelements; ++eM) {
    for (int nz = 0; nz < Nq; nz++)
      for (int ny = 0; ny < Nq; ny++)
        for (int nx = 0; nx < Nq; nx++) {
          int idn = nx + ny * Nq + nz * Nq * Nq;
          double& a = A[(size_t)eM * Np + idn];
          if (mapB[idn + (size_t)eM * Np] != 1) {
            a = 0;
            const size_t base = (size_t)eM * Np * 6;
            a += 2 * mesh.ggeo[base + idn + 1 * Np] * D[nx + nx * Nq] * D[ny + ny * Nq];
            a += 2 * mesh.ggeo[base + idn + 2 * Np] * D[nx + nx * Nq] * D[nz + nz * Nq];
            a += 2 * mesh.ggeo[base + idn + 4 * Np] * D[ny + ny * Nq] * D[nz + nz * Nq];
            for (int k = 0; k < Nq; k++) {
              int iid = k + ny * Nq + nz * Nq * Nq;
              a += mesh.ggeo[base + iid + 0 * Np] * D[nx + k * Nq] * D[nx + k * Nq];
            }
            for (int k = 0; k < Nq; k++) {
              int iid = nx + k * Nq + nz * Nq * Nq;
              a += mesh.ggeo[base + iid + 3 * Np] * D[ny + k * Nq] * D[ny + k * Nq];
            }
            for (int k = 0; k < Nq; k++) {
              int iid = nx + ny * Nq + k * Nq * Nq;
              a += mesh.ggeo[base + iid + 5 * Np] * D[nz + k * Nq] * D[nz + k * Nq];
            }
            a += mesh.wJ[(size_t)eM * Np + idn] * lambda;
          } else {
            a = 1;
          }
        }
    if (allNeumann)
      for (int n = 0; n < Np; ++n)
        if (mapB[n + (size_t)eM * Np] != 1) A[(size_t)eM * Np + n] += allNeumannPenalty * allNeumannScale * allNeumannScale;
  }
  diagA.assign(ogsMasked.Ngather, 0.0);
  ogsMasked.gatherLocal.GatherAdd(diagA.data(), A.data(), true);
}

/* ellipticZeroMean.cpp:28-36. */
void Elliptic::ZeroMean(double* q) const {
  double qmean = linalg_sum(Ndofs, q);
  qmean *= allNeumannScale * allNeumannScale;
  for (dlong n = 0; n < Ndofs; ++n) q[n] += -qmean;
}

/* Right-hand side in local (E-vector) form: the forcing kernel
 * (okl/ellipticRhsHex3D.okl:28-51) followed by the C0 boundary kernel
 * (okl/ellipticRhsBCHex3D.okl:28-289). For Dirichlet (bc==1) faces the flux
 * term vanishes (the kernel passes zero normal derivatives), so the boundary
 * kernel reduces to rhs -= A_L u_b with u_b the Dirichlet data on the faces
 * whose node flag is set. Neumann faces (bc==2) are not used by the box
 * configurations and are rejected. */
void Elliptic::Rhs(std::vector<double>& rL) const {
  DataFn data;
  data.setup(settings);
  const int Nq = mesh.Nq, Np = mesh.Np, NqNq = Nq * Nq;
  const size_t NL = (size_t)mesh.Nelements * Np;
  rL.assign(NL, 0.0);
  for (size_t id = 0; id < NL; ++id)
    rL[id] = mesh.wJ[id] * data.forcing(mesh.x[id], mesh.y[id], mesh.z[id], lambda);
  for (size_t id = 0; id < NL; ++id)
    if (mapB[id] == 2) throw std::runtime_error("Neumann boundary data not supported by the oracle");
  // u_b on the element's face nodes (zero elsewhere); the face loops of the
  // kernel visit only face nodes and use mapB of that node.
  std::vector<double> ub(Np), Aub(Np);
  const double* DT = mesh.D.data();
  for (dlong e = 0; e < mesh.Nelements; ++e) {
    bool any = false;
    for (int n = 0; n < Np; ++n) ub[n] = 0;
    for (int f = 0; f < 6; ++f)
      for (int m = 0; m < mesh.Nfp; ++m) {
        int n = mesh.faceNodes[f * mesh.Nfp + m];
        size_t id = (size_t)e * Np + n;
        if (mapB[id] > 0) {
          ub[n] = (mapB[id] == 1) ? data.dirichlet(mesh.x[id], mesh.y[id], mesh.z[id]) : 0.0;
          any = true;
        }
      }
    if (!any) continue;
    // Same sweep structure as the partial Ax on the element-local u_b.
    std::vector<double> r_q(Np), r_rhs(Np, 0.0), s_q(NqNq), s0(NqNq), s1(NqNq), qtv(NqNq), Gqt(NqNq), Auk(NqNq);
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i)
        for (int k = 0; k < Nq; ++k) r_q[(j * Nq + i) * Nq + k] = ub[i + j * Nq + k * NqNq];
    for (int k = 0; k < Nq; ++k) {
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const double* rq = &r_q[(j * Nq + i) * Nq];
          s_q[j * Nq + i] = rq[k];
          double qt = 0;
          for (int m = 0; m < Nq; ++m) qt += DT[k * Nq + m] * rq[m];
          qtv[j * Nq + i] = qt;
        }
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          double qr = 0, qs = 0;
          for (int m = 0; m < Nq; ++m) {
            qr += DT[i * Nq + m] * s_q[j * Nq + m];
            qs += DT[j * Nq + m] * s_q[m * Nq + i];
          }
          const size_t gb = (size_t)e * 6 * Np + k * NqNq + j * Nq + i;
          const double G00 = mesh.ggeo[gb], G01 = mesh.ggeo[gb + Np], G02 = mesh.ggeo[gb + 2 * Np];
          const double G11 = mesh.ggeo[gb + 3 * Np], G12 = mesh.ggeo[gb + 4 * Np], G22 = mesh.ggeo[gb + 5 * Np];
          const double GwJ = mesh.wJ[(size_t)e * Np + k * NqNq + j * Nq + i];
          const double qt = qtv[j * Nq + i];
          s0[j * Nq + i] = (G01 * qr + G11 * qs + G12 * qt);
          s1[j * Nq + i] = (G00 * qr + G01 * qs + G02 * qt);
          Gqt[j * Nq + i] = (G02 * qr + G12 * qs + G22 * qt);
          Auk[j * Nq + i] = GwJ * lambda * r_q[(j * Nq + i) * Nq + k];
        }
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          double a = Auk[j * Nq + i];
          double* rr = &r_rhs[(j * Nq + i) * Nq];
          for (int m = 0; m < Nq; ++m) {
            a += DT[m * Nq + j] * s0[m * Nq + i];
            rr[m] += DT[k * Nq + m] * Gqt[j * Nq + i];
            a += DT[m * Nq + i] * s1[j * Nq + m];
          }
          rr[k] += a;
        }
    }
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i)
        for (int k = 0; k < Nq; ++k) rL[(size_t)e * Np + k * NqNq + j * Nq + i] -= r_rhs[(j * Nq + i) * Nq + k];
  }
}

/* okl/ellipticAddBCHex3D.okl:27-47. */
void Elliptic::AddBC(std::vector<double>& xL) const {
  DataFn data;
  data.setup(settings);
  const size_t NL = (size_t)mesh.Nelements * mesh.Np;
  for (size_t id = 0; id < NL; ++id)
    if (mapB[id] == 1) xL[id] = data.dirichlet(mesh.x[id], mesh.y[id], mesh.z[id]);
}

}  // namespace oracle
