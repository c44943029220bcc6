// CPU oracle (test infrastructure only): hex box mesh, connectivity, node
// numbering and volume geometric factors, restated from libs/mesh/*.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

static const double PI_ = 3.14159265358979323846;

/* meshSetupBoxHex3D.cpp:31-232 (size=1, rank=0: offsets are zero). */
void Mesh::SetupBox(const Settings& st) {
  dlong NX = st.geti("BOX GLOBAL NX"), NY = st.geti("BOX GLOBAL NY"), NZ = st.geti("BOX GLOBAL NZ");
  nx = st.geti("BOX NX");
  ny = st.geti("BOX NY");
  nz = st.geti("BOX NZ");
  if ((long long)NX * NY * NZ <= 0) {
    NX = nx; NY = ny; NZ = nz;
  } else {
    nx = NX; ny = NY; nz = NZ;
  }
  boundaryFlag = st.geti("BOX BOUNDARY FLAG");
  const int periodicFlag = (boundaryFlag == -1) ? 1 : 0;
  DIMX = st.getd("BOX DIMX");
  DIMY = st.getd("BOX DIMY");
  DIMZ = st.getd("BOX DIMZ");
  if (st.has("BOX CURVED MAP AMPLITUDE")) curveAmp = st.getd("BOX CURVED MAP AMPLITUDE");

  const double dx = DIMX / NX, dy = DIMY / NY, dz = DIMZ / NZ;
  const double X0 = -DIMX / 2.0, Y0 = -DIMY / 2.0, Z0 = -DIMZ / 2.0;
  const hlong NnX = periodicFlag ? NX : NX + 1;
  const hlong NnY = periodicFlag ? NY : NY + 1;
  const hlong NnZ = periodicFlag ? NZ : NZ + 1;
  Nnodes = NnX * NnY * NnZ;
  Nelements = nx * ny * nz;
  EToV.assign((size_t)Nelements * Nverts, 0);
  EX.assign((size_t)Nelements * Nverts, 0);
  EY.assign((size_t)Nelements * Nverts, 0);
  EZ.assign((size_t)Nelements * Nverts, 0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const dlong e = i + j * nx + k * nx * ny;
        const hlong i0 = i, i1 = (i + 1) % NnX, j0 = j, j1 = (j + 1) % NnY, k0 = k, k1 = (k + 1) % NnZ;
        hlong* v = &EToV[(size_t)e * Nverts];
        v[0] = i0 + j0 * NnX + k0 * NnX * NnY;
        v[1] = i1 + j0 * NnX + k0 * NnX * NnY;
        v[2] = i1 + j1 * NnX + k0 * NnX * NnY;
        v[3] = i0 + j1 * NnX + k0 * NnX * NnY;
        v[4] = i0 + j0 * NnX + k1 * NnX * NnY;
        v[5] = i1 + j0 * NnX + k1 * NnX * NnY;
        v[6] = i1 + j1 * NnX + k1 * NnX * NnY;
        v[7] = i0 + j1 * NnX + k1 * NnX * NnY;
        double x0 = X0 + dx * i, y0 = Y0 + dy * j, z0 = Z0 + dz * k;
        double* ex = &EX[(size_t)e * Nverts];
        double* ey = &EY[(size_t)e * Nverts];
        double* ez = &EZ[(size_t)e * Nverts];
        ex[0] = x0;      ey[0] = y0;      ez[0] = z0;
        ex[1] = x0 + dx; ey[1] = y0;      ez[1] = z0;
        ex[2] = x0 + dx; ey[2] = y0 + dy; ez[2] = z0;
        ex[3] = x0;      ey[3] = y0 + dy; ez[3] = z0;
        ex[4] = x0;      ey[4] = y0;      ez[4] = z0 + dz;
        ex[5] = x0 + dx; ey[5] = y0;      ez[5] = z0 + dz;
        ex[6] = x0 + dx; ey[6] = y0 + dy; ez[6] = z0 + dz;
        ex[7] = x0;      ey[7] = y0 + dy; ez[7] = z0 + dz;
      }
  boundaryInfo.clear();
  if (boundaryFlag != -1) {
    NboundaryFaces = 2 * NX * NY + 2 * NX * NZ + 2 * NY * NZ;
    auto add = [&](hlong a, hlong b, hlong c, hlong d) {
      boundaryInfo.push_back(boundaryFlag);
      boundaryInfo.push_back(a);
      boundaryInfo.push_back(b);
      boundaryInfo.push_back(c);
      boundaryInfo.push_back(d);
    };
    for (hlong j = 0; j < NY; ++j)
      for (hlong i = 0; i < NX; ++i) {
        hlong vid1 = i + j * NnX, vid2 = i + j * NnX + NZ * NnX * NnY;
        add(vid1, vid1 + 1, vid1 + 1 + NnX, vid1 + NnX);
        add(vid2, vid2 + 1, vid2 + 1 + NnX, vid2 + NnX);
      }
    for (hlong k = 0; k < NZ; ++k)
      for (hlong i = 0; i < NX; ++i) {
        hlong vid1 = i + k * NnX * NnY, vid2 = i + NY * NnX + k * NnX * NnY;
        add(vid1, vid1 + 1, vid1 + 1 + NnX * NnY, vid1 + NnX * NnY);
        add(vid2, vid2 + 1, vid2 + 1 + NnX * NnY, vid2 + NnX * NnY);
      }
    for (hlong k = 0; k < NZ; ++k)
      for (hlong j = 0; j < NY; ++j) {
        hlong vid1 = j * NnX + k * NnX * NnY, vid2 = NX + j * NnX + k * NnX * NnY;
        add(vid1, vid1 + NnX, vid1 + NnX + NnX * NnY, vid1 + NnX * NnY);
        add(vid2, vid2 + NnX, vid2 + NnX + NnX * NnY, vid2 + NnX * NnY);
      }
  } else {
    NboundaryFaces = 0;
  }
  // meshSetElementType.cpp (HEXAHEDRA): vertices on each face
  static const int fv[6][4] = {{0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3}, {0, 3, 7, 4}, {4, 7, 6, 5}};
  faceVertices.assign(&fv[0][0], &fv[0][0] + 24);
}

/* meshReferenceNodesHex3D.cpp:31-74 and meshBasisHex3D.cpp:34-93. */
void Mesh::ReferenceNodes(int N_) {
  N = N_;
  Nq = N + 1;
  Nfp = Nq * Nq;
  Np = Nq * Nq * Nq;
  JacobiGLL(N, gllz, gllw);
  r.resize(Np); s.resize(Np); t.resize(Np);
  for (int k = 0; k < Nq; k++)
    for (int j = 0; j < Nq; j++)
      for (int i = 0; i < Nq; i++) {
        r[i + j * Nq + k * Nq * Nq] = gllz[i];
        s[i + j * Nq + k * Nq * Nq] = gllz[j];
        t[i + j * Nq + k * Nq * Nq] = gllz[k];
      }
  double deps = 1.;
  while ((1. + deps) > 1.) deps *= 0.5;
  const double NODETOL = 1000. * deps;
  faceNodes.assign(6 * Nfp, 0);
  int cnt[6] = {0, 0, 0, 0, 0, 0};
  for (int n = 0; n < Np; n++) {
    if (std::fabs(t[n] + 1) < NODETOL) faceNodes[0 * Nfp + (cnt[0]++)] = n;
    if (std::fabs(s[n] + 1) < NODETOL) faceNodes[1 * Nfp + (cnt[1]++)] = n;
    if (std::fabs(r[n] - 1) < NODETOL) faceNodes[2 * Nfp + (cnt[2]++)] = n;
    if (std::fabs(s[n] - 1) < NODETOL) faceNodes[3 * Nfp + (cnt[3]++)] = n;
    if (std::fabs(r[n] + 1) < NODETOL) faceNodes[4 * Nfp + (cnt[4]++)] = n;
    if (std::fabs(t[n] - 1) < NODETOL) faceNodes[5 * Nfp + (cnt[5]++)] = n;
  }
  Dmatrix1D(N, gllz, D);
}

/* meshConnect.cpp:51-130 (local part; the interprocess part is empty on 1 rank). */
namespace {
struct Face {
  hlong v[4];
  dlong element, elementN;
  int face, faceN;
};
}  // namespace

void Mesh::Connect() {
  const dlong NF = Nelements * Nfaces;
  std::vector<Face> faces(NF);
  for (dlong e = 0; e < Nelements; ++e)
    for (int f = 0; f < Nfaces; ++f) {
      Face& F = faces[f + e * Nfaces];
      for (int n = 0; n < NfaceVertices; ++n) F.v[n] = EToV[(size_t)e * Nverts + faceVertices[f * NfaceVertices + n]];
      std::sort(F.v, F.v + NfaceVertices);
      F.element = e;
      F.face = f;
      F.elementN = -1;
      F.faceN = -1;
    }
  std::sort(faces.begin(), faces.end(), [](const Face& a, const Face& b) {
    return std::lexicographical_compare(a.v, a.v + 4, b.v, b.v + 4);
  });
  for (dlong c = 0; c + 1 < NF; ++c) {
    if (std::equal(faces[c].v, faces[c].v + 4, faces[c + 1].v)) {
      faces[c].elementN = faces[c + 1].element;
      faces[c].faceN = faces[c + 1].face;
      faces[c + 1].elementN = faces[c].element;
      faces[c + 1].faceN = faces[c].face;
    }
  }
  std::sort(faces.begin(), faces.end(), [](const Face& a, const Face& b) {
    if (a.element < b.element) return true;
    if (a.element > b.element) return false;
    return a.face < b.face;
  });
  EToE.assign(NF, -1);
  EToF.assign(NF, -1);
  for (dlong id = 0; id < NF; ++id) {
    EToE[id] = faces[id].elementN;
    EToF[id] = faces[id].faceN;
  }
}

/* meshConnectBoundary.cpp:54-156. */
void Mesh::ConnectBoundary() {
  struct BF {
    dlong element;
    int face;
    hlong v[4];
    int bctype;
  };
  std::vector<BF> bfs;
  for (dlong e = 0; e < Nelements; ++e)
    for (int f = 0; f < Nfaces; ++f)
      if (EToE[e * Nfaces + f] == -1) {
        BF b;
        for (int n = 0; n < NfaceVertices; ++n) b.v[n] = EToV[(size_t)e * Nverts + faceVertices[f * NfaceVertices + n]];
        std::sort(b.v, b.v + 4);
        b.element = e;
        b.face = f;
        b.bctype = -1;
        bfs.push_back(b);
      }
  for (hlong bb = 0; bb < NboundaryFaces; ++bb) {
    BF b;
    for (int n = 0; n < NfaceVertices; ++n) b.v[n] = boundaryInfo[bb * (NfaceVertices + 1) + n + 1];
    std::sort(b.v, b.v + 4);
    b.element = -1;
    b.face = -1;
    b.bctype = (int)boundaryInfo[bb * (NfaceVertices + 1)];
    bfs.push_back(b);
  }
  std::sort(bfs.begin(), bfs.end(), [](const BF& a, const BF& b) {
    return std::lexicographical_compare(a.v, a.v + 4, b.v, b.v + 4);
  });
  EToB.assign((size_t)Nelements * Nfaces, -1);
  for (size_t c = 0; c + 1 < bfs.size(); ++c) {
    if (std::equal(bfs[c].v, bfs[c].v + 4, bfs[c + 1].v)) {
      dlong e = std::max(bfs[c].element, bfs[c + 1].element);
      int f = std::max(bfs[c].face, bfs[c + 1].face);
      EToB[e * Nfaces + f] = std::max(bfs[c].bctype, bfs[c + 1].bctype);
    }
  }
}

/* FaceNodeMatchingHex3D (meshBasisHex3D.cpp:131-263) + meshConnectFaceNodes.cpp:32-101. */
void Mesh::ConnectFaceNodes() {
  const double NODETOL = 1.0e-5;
  const double V0[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
  const double V1[4][2] = {{-1.0, -1.0}, {-1.0, 1.0}, {1.0, 1.0}, {1.0, -1.0}};
  std::vector<int> R(Nfaces * Nfaces * NfaceVertices * Nfp, 0);
  auto tri = [&](const double* EXv, int fn) {
    double rn = r[fn], sn = s[fn], tn = t[fn];
    return +0.125 * (1 - rn) * (1 - sn) * (1 - tn) * EXv[0] + 0.125 * (1 + rn) * (1 - sn) * (1 - tn) * EXv[1] +
           0.125 * (1 + rn) * (1 + sn) * (1 - tn) * EXv[2] + 0.125 * (1 - rn) * (1 + sn) * (1 - tn) * EXv[3] +
           0.125 * (1 - rn) * (1 - sn) * (1 + tn) * EXv[4] + 0.125 * (1 + rn) * (1 - sn) * (1 + tn) * EXv[5] +
           0.125 * (1 + rn) * (1 + sn) * (1 + tn) * EXv[6] + 0.125 * (1 - rn) * (1 + sn) * (1 + tn) * EXv[7];
  };
  std::vector<double> x0(Nfp), y0(Nfp), x1(Nfp), y1(Nfp);
  for (int fM = 0; fM < Nfaces; fM++) {
    double EX0[8] = {0}, EY0[8] = {0};
    for (int v = 0; v < NfaceVertices; v++) {
      int fv = faceVertices[fM * NfaceVertices + v];
      EX0[fv] = V0[v][0];
      EY0[fv] = V0[v][1];
    }
    for (int n = 0; n < Nfp; ++n) {
      const int fn = faceNodes[fM * Nfp + n];
      x0[n] = tri(EX0, fn);
      y0[n] = tri(EY0, fn);
    }
    for (int fP = 0; fP < Nfaces; fP++)
      for (int rot = 0; rot < NfaceVertices; rot++) {
        double EX1[8] = {0}, EY1[8] = {0};
        for (int v = 0; v < NfaceVertices; v++) {
          int fv = faceVertices[fP * NfaceVertices + ((v + rot) % NfaceVertices)];
          EX1[fv] = V1[v][0];
          EY1[fv] = V1[v][1];
        }
        for (int n = 0; n < Nfp; ++n) {
          const int fn = faceNodes[fP * Nfp + n];
          x1[n] = tri(EX1, fn);
          y1[n] = tri(EY1, fn);
        }
        for (int n = 0; n < Nfp; ++n) {
          int m = 0;
          for (; m < Nfp; ++m) {
            double dist = std::pow(x0[n] - x1[m], 2) + std::pow(y0[n] - y1[m], 2);
            if (dist < NODETOL) {
              R[fM * Nfaces * NfaceVertices * Nfp + fP * NfaceVertices * Nfp + rot * Nfp + n] = m;
              break;
            }
          }
          if (m == Nfp) throw std::runtime_error("Unable to match face node");
        }
      }
  }
  const size_t NT = (size_t)Nfp * Nfaces * Nelements;
  vmapM.assign(NT, 0);
  vmapP.assign(NT, 0);
  for (dlong eM = 0; eM < Nelements; ++eM)
    for (int fM = 0; fM < Nfaces; ++fM) {
      dlong eP = EToE[eM * Nfaces + fM];
      int fP = EToF[eM * Nfaces + fM];
      if (eP < 0 || fP < 0) {
        for (int nM = 0; nM < Nfp; ++nM) {
          const int idM = faceNodes[fM * Nfp + nM];
          const size_t id = (size_t)eM * Nfaces * Nfp + fM * Nfp + nM;
          vmapM[id] = idM + eM * Np;
          vmapP[id] = idM + eM * Np;
        }
      } else {
        hlong vf0P = EToV[(size_t)eP * Nverts + faceVertices[fP * NfaceVertices + 0]];
        int rot = 0;
        for (; rot < NfaceVertices; ++rot) {
          hlong vfM = EToV[(size_t)eM * Nverts + faceVertices[fM * NfaceVertices + rot]];
          if (vfM == vf0P) break;
        }
        for (int nM = 0; nM < Nfp; ++nM) {
          const int nP = R[fM * Nfaces * Nfp * NfaceVertices + fP * Nfp * NfaceVertices + rot * Nfp + nM];
          const int idM = faceNodes[fM * Nfp + nM];
          const int idP = faceNodes[fP * Nfp + nP];
          const size_t id = (size_t)eM * Nfaces * Nfp + fM * Nfp + nM;
          vmapM[id] = idM + eM * Np;
          vmapP[id] = idP + eP * Np;
        }
      }
    }
}

/* meshConnectNodes.cpp:32-117 (gatherNodeStart = 0 on one rank). */
void Mesh::ConnectNodes() {
  const size_t NL = (size_t)Nelements * Np;
  globalIds.resize(NL);
  for (size_t n = 0; n < NL; ++n) globalIds[n] = 1 + (hlong)n;
  mapB.assign(NL, -1);
  for (dlong e = 0; e < Nelements; e++)
    for (int f = 0; f < Nfaces; f++) {
      int bc = EToB[f + e * Nfaces];
      if (bc > 0) {
        for (int n = 0; n < Nfp; n++) {
          const int fid = faceNodes[n + f * Nfp];
          int bcn = mapB[fid + (size_t)e * Np];
          mapB[fid + (size_t)e * Np] = (bcn == -1) ? bc : std::min(bc, bcn);
        }
      }
    }
  hlong gatherChange = 1;
  while (gatherChange > 0) {
    gatherChange = 0;
    for (dlong e = 0; e < Nelements; ++e)
      for (int n = 0; n < Nfp * Nfaces; ++n) {
        size_t id = (size_t)e * Nfp * Nfaces + n;
        dlong idM = vmapM[id], idP = vmapP[id];
        hlong gidM = globalIds[idM], gidP = globalIds[idP];
        int bcM = mapB[idM], bcP = mapB[idP];
        if (gidP < gidM) {
          ++gatherChange;
          globalIds[idM] = gidP;
        }
        if (bcP > 0) {
          if (bcM == -1) {
            mapB[idM] = bcP;
            ++gatherChange;
          } else if (bcP < bcM) {
            mapB[idM] = bcP;
            ++gatherChange;
          }
        }
      }
  }
}

/* meshPhysicalNodesHex3D.cpp:31-112, followed by the optional curved map
 * x <- x + A*sin(2 pi y)*sin(2 pi z) (no reference counterpart). */
void Mesh::PhysicalNodes() {
  const size_t NL = (size_t)Nelements * Np;
  x.resize(NL); y.resize(NL); z.resize(NL);
  for (dlong e = 0; e < Nelements; ++e) {
    const double* xe = &EX[(size_t)e * Nverts];
    const double* ye = &EY[(size_t)e * Nverts];
    const double* ze = &EZ[(size_t)e * Nverts];
    for (int n = 0; n < Np; ++n) {
      double rn = r[n], sn = s[n], tn = t[n];
      double c[8] = {0.125 * (1 - rn) * (1 - sn) * (1 - tn), 0.125 * (1 + rn) * (1 - sn) * (1 - tn),
                     0.125 * (1 + rn) * (1 + sn) * (1 - tn), 0.125 * (1 - rn) * (1 + sn) * (1 - tn),
                     0.125 * (1 - rn) * (1 - sn) * (1 + tn), 0.125 * (1 + rn) * (1 - sn) * (1 + tn),
                     0.125 * (1 + rn) * (1 + sn) * (1 + tn), 0.125 * (1 - rn) * (1 + sn) * (1 + tn)};
      double X = 0, Y = 0, Z = 0;
      for (int v = 0; v < 8; ++v) {
        X += c[v] * xe[v];
        Y += c[v] * ye[v];
        Z += c[v] * ze[v];
      }
      if (curveAmp != 0.0) X += curveAmp * std::sin(2 * PI_ * Y) * std::sin(2 * PI_ * Z);
      x[(size_t)e * Np + n] = X;
      y[(size_t)e * Np + n] = Y;
      z[(size_t)e * Np + n] = Z;
    }
  }
}

/* meshGeometricFactorsHex3D.cpp:31-193: ggeo (6 per node) and wJ only;
 * vgeo is not used by the C0 path. */
void Mesh::GeometricFactors() {
  const size_t NL = (size_t)Nelements * Np;
  ggeo.assign(NL * 6, 0.0);
  wJ.assign(NL, 0.0);
  for (dlong e = 0; e < Nelements; ++e)
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          int n = i + j * Nq + k * Nq * Nq;
          double xr = 0, xs = 0, xt = 0, yr = 0, ys = 0, yt = 0, zr = 0, zs = 0, zt = 0;
          for (int m = 0; m < Nq; ++m) {
            size_t idr = (size_t)e * Np + k * Nq * Nq + j * Nq + m;
            size_t ids = (size_t)e * Np + k * Nq * Nq + m * Nq + i;
            size_t idt = (size_t)e * Np + m * Nq * Nq + j * Nq + i;
            xr += D[i * Nq + m] * x[idr]; xs += D[j * Nq + m] * x[ids]; xt += D[k * Nq + m] * x[idt];
            yr += D[i * Nq + m] * y[idr]; ys += D[j * Nq + m] * y[ids]; yt += D[k * Nq + m] * y[idt];
            zr += D[i * Nq + m] * z[idr]; zs += D[j * Nq + m] * z[ids]; zt += D[k * Nq + m] * z[idt];
          }
          double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
          if (J < 1e-12) throw std::runtime_error("Negative J found at element " + std::to_string(e));
          double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
          double sx = -(yr * zt - zr * yt) / J, sy = (xr * zt - zr * xt) / J, sz = -(xr * yt - yr * xt) / J;
          double tx = (yr * zs - zr * ys) / J, ty = -(xr * zs - zr * xs) / J, tz = (xr * ys - yr * xs) / J;
          double JW = J * gllw[i] * gllw[j] * gllw[k];
          size_t b = (size_t)6 * Np * e + n;
          ggeo[b + Np * 0] = JW * (rx * rx + ry * ry + rz * rz);
          ggeo[b + Np * 1] = JW * (rx * sx + ry * sy + rz * sz);
          ggeo[b + Np * 2] = JW * (rx * tx + ry * ty + rz * tz);
          ggeo[b + Np * 3] = JW * (sx * sx + sy * sy + sz * sz);
          ggeo[b + Np * 4] = JW * (sx * tx + sy * ty + sz * tz);
          ggeo[b + Np * 5] = JW * (tx * tx + ty * ty + tz * tz);
          wJ[(size_t)Np * e + n] = JW;
        }
}

/* meshGatherScatterSetup.cpp:32-130: on one rank no element touches a halo
 * vertex, so every element is a "local gather" element, in order. */
void Mesh::GatherScatterSetup() {
  localGatherElementList.resize(Nelements);
  for (dlong e = 0; e < Nelements; ++e) localGatherElementList[e] = e;
}

/* meshSetup.cpp:31-102 (BOX branch). */
void Mesh::Setup(const Settings& st) {
  SetupBox(st);
  ReferenceNodes(st.geti("POLYNOMIAL DEGREE"));
  Connect();
  ConnectBoundary();
  ConnectFaceNodes();
  ConnectNodes();
  PhysicalNodes();
  GeometricFactors();
  GatherScatterSetup();
}

/* meshSetupNewDegree.cpp:32-64. */
Mesh Mesh::SetupNewDegree(int Nf) const {
  Mesh m = *this;
  if (Nf == N) return m;
  m.ReferenceNodes(Nf);
  m.ConnectFaceNodes();
  m.ConnectNodes();
  m.PhysicalNodes();
  m.GeometricFactors();
  m.GatherScatterSetup();
  return m;
}

}  // namespace oracle
