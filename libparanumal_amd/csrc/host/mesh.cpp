// Hex meshes: a box (structured connectivity in closed form) or a Gmsh file
// (general numbering); geometry on the GPU.
#include "mesh.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>

namespace libp {

static void legendre(int N, double x, double& P, double& dP) {
  // P_N(x) and P_N'(x) by the three-term recurrence
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
  if (N == 0) {
    P = 1.0;
    dP = 0.0;
    return;
  }
  for (int n = 1; n < N; ++n) {
    const double p2 = ((2 * n + 1) * x * p1 - n * p0) / (n + 1);
    const double d2 = d0 + (2 * n + 1) * p1;
    p0 = p1;
    p1 = p2;
    d0 = d1;
    d1 = d2;
  }
  P = p1;
  dP = d1;
}

void gllNodesWeights(int N, std::vector<double>& r, std::vector<double>& w) {
  r.assign(N + 1, 0.0);
  w.assign(N + 1, 0.0);
  r[0] = -1.0;
  r[N] = 1.0;
  // interior nodes: roots of P_N'; Newton on q(x) = P_N'(x) with
  // q'(x) = (2x P_N' - N(N+1) P_N)/(1-x^2), from Chebyshev-Lobatto guesses.
  for (int i = 1; 2 * i <= N; ++i) {
    double x = -std::cos(M_PI * i / N);
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre(N, x, P, dP);
      const double ddP = (2 * x * dP - N * (N + 1) * P) / (1 - x * x);
      const double dx = dP / ddP;
      x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    r[i] = x;
    r[N - i] = -x;
  }
  if (N % 2 == 0) r[N / 2] = 0.0;
  for (int i = 0; i <= N; ++i) {
    double P, dP;
    legendre(N, r[i], P, dP);
    w[i] = 2.0 / (N * (N + 1) * P * P);
  }
}

void gllDmatrix(int N, const std::vector<double>& r, std::vector<double>& D) {
  const int Nq = N + 1;
  D.assign(Nq * Nq, 0.0);
  if (N == 0) return;
  std::vector<double> P(Nq);
  for (int i = 0; i < Nq; ++i) {
    double dP;
    legendre(N, r[i], P[i], dP);
  }
  for (int i = 0; i < Nq; ++i)
    for (int j = 0; j < Nq; ++j) {
      if (i != j) D[i * Nq + j] = P[i] / (P[j] * (r[i] - r[j]));
    }
  D[0] = -0.25 * N * (N + 1);
  D[N * Nq + N] = 0.25 * N * (N + 1);
}

void degreeRaiseMatrix(int Nc, int Nf, std::vector<double>& P) {
  std::vector<double> rc, wc, rf, wf;
  gllNodesWeights(Nc, rc, wc);
  gllNodesWeights(Nf, rf, wf);
  const int NqC = Nc + 1, NqF = Nf + 1;
  P.assign(NqF * NqC, 0.0);
  for (int f = 0; f < NqF; ++f)
    for (int c = 0; c < NqC; ++c) {
      double l = 1.0;
      for (int k = 0; k < NqC; ++k)
        if (k != c) l *= (rf[f] - rc[k]) / (rc[c] - rc[k]);
      P[f * NqC + c] = l;
    }
}

void mesh_t::Setup(platform_t& _platform, settings_t& _settings, comm_t _comm) {
  platform = _platform;
  settings = _settings;
  comm = _comm;
  LIBP_ABORT("Only hexahedral meshes (ELEMENT TYPE 12) are supported",
             !settings.compareSetting("ELEMENT TYPE", "12"));
  LIBP_ABORT("Only 3-D meshes are supported", !settings.compareSetting("MESH DIMENSION", "3"));
  const std::string file = settings.getSetting("MESH FILE");
  boxMesh = file == "BOX";
  if (boxMesh)
    SetupBox();
  else
    ReadGmshHex3D(file);
  int NN;
  settings.getSetting("POLYNOMIAL DEGREE", NN);
  ReferenceNodes(NN);
  if (boxMesh) {
    ConnectNodes();
    GeometricFactors();
  } else {
    GeometricFactors();  // the general numbering matches nodes by their coordinates
    ConnectNodesGeneral();
  }
  GatherScatterSetup();
}

/* Gmsh 2.2 ASCII hexahedral mesh (meshReadGmshHex3D.cpp): $Nodes "id x y z";
 * $Elements "id type ntags tag... nodes", type 5 = hexahedron (8 vertices, in
 * the order of the reference hex), type 3 = quadrilateral boundary face
 * whose first tag is the boundary flag. One rank reads every element. */
void mesh_t::ReadGmshHex3D(const std::string& fileName) {
  std::ifstream f(fileName);
  LIBP_ABORT("Cannot open file: " << fileName, !f);
  std::string line;
  auto seek = [&](const char* tag) {
    while (std::getline(f, line))
      if (line.find(tag) != std::string::npos) return;
    LIBP_FORCE_ABORT("Error reading mesh file: " << fileName << " (no " << tag << ")");
  };
  seek("$Nodes");
  hlong Nnodes = 0;
  LIBP_ABORT("Error reading mesh file: " << fileName, !std::getline(f, line));
  Nnodes = std::stoll(line);
  std::vector<double> VX(Nnodes), VY(Nnodes), VZ(Nnodes);
  for (hlong n = 0; n < Nnodes; ++n) {
    LIBP_ABORT("Error reading mesh file: " << fileName, !std::getline(f, line));
    std::istringstream ss(line);
    hlong id;
    ss >> id >> VX[n] >> VY[n] >> VZ[n];
    LIBP_ABORT("Error reading node " << n + 1 << " of " << fileName, !ss);
  }
  seek("$Elements");
  LIBP_ABORT("Error reading mesh file: " << fileName, !std::getline(f, line));
  const hlong gN = std::stoll(line);
  std::vector<hlong> hexes;
  boundaryInfo.clear();
  for (hlong n = 0; n < gN; ++n) {
    LIBP_ABORT("Error reading mesh file: " << fileName, !std::getline(f, line));
    std::istringstream ss(line);
    hlong id;
    int type, ntags;
    ss >> id >> type >> ntags;
    std::vector<hlong> tags(ntags);
    for (auto& t : tags) ss >> t;
    if (type == 5) {
      for (int v = 0; v < 8; ++v) {
        hlong vid;
        ss >> vid;
        LIBP_ABORT("vertex id " << vid << " out of range in " << fileName, vid < 1 || vid > Nnodes);
        hexes.push_back(vid - 1);
      }
    } else if (type == 3) {
      boundaryInfo.push_back(ntags > 0 ? tags[0] : 0);
      for (int v = 0; v < 4; ++v) {
        hlong vid;
        ss >> vid;
        LIBP_ABORT("vertex id " << vid << " out of range in " << fileName, vid < 1 || vid > Nnodes);
        boundaryInfo.push_back(vid - 1);
      }
    }
    LIBP_ABORT("Error reading element " << id << " of " << fileName, !ss);
  }
  Nelements = (dlong)(hexes.size() / 8);
  LIBP_ABORT("no hexahedra in " << fileName, Nelements == 0);
  nx = ny = nz = 0;
  periodic = false;
  boundaryFlag = 0;
  curveAmp = 0.0;
  EToV.malloc((size_t)Nelements * Nverts);
  EX.malloc((size_t)Nelements * Nverts);
  EY.malloc((size_t)Nelements * Nverts);
  EZ.malloc((size_t)Nelements * Nverts);
  for (size_t k = 0; k < hexes.size(); ++k) {
    EToV[k] = hexes[k];
    EX[k] = VX[hexes[k]];
    EY[k] = VY[hexes[k]];
    EZ[k] = VZ[hexes[k]];
  }
  o_EX = platform.malloc(EX);
  o_EY = platform.malloc(EY);
  o_EZ = platform.malloc(EZ);
}

/* Node numbering of a mesh file. On one rank the reference's fixed point
 * (meshConnectNodes.cpp:32-117) gives every node 1 + the smallest local
 * index among its copies; here the copies are found by their coordinates
 * (hashed on a grid of 1e-6 of the smallest element extent, the 27
 * neighbouring cells searched, copies within 1e-9 of it). mapB: the smallest
 * positive flag of the boundary faces (the mesh file's tagged quadrilaterals,
 * matched by vertex set as meshConnectBoundary.cpp:54-156 does) touching any
 * copy of the node, else -1. */
void mesh_t::ConnectNodesGeneral() {
  const size_t NL = (size_t)Nelements * Np;
  std::vector<double> x(NL), y(NL), z(NL);
  o_x.copyTo(x.data(), NL, 0, platform.stream);
  o_y.copyTo(y.data(), NL, 0, platform.stream);
  o_z.copyTo(z.data(), NL, 0, platform.stream);
  platform.finish();
  double h = 1e300;  // smallest element extent along any axis
  for (dlong e = 0; e < Nelements; ++e)
    for (const memory<double>* V : {&EX, &EY, &EZ}) {
      double lo = 1e300, hi = -1e300;
      for (int v = 0; v < 8; ++v) {
        lo = std::min(lo, (*V)[(size_t)e * 8 + v]);
        hi = std::max(hi, (*V)[(size_t)e * 8 + v]);
      }
      h = std::min(h, hi - lo);
    }
  LIBP_ABORT("degenerate element extent", !(h > 0));
  const double cell = 1e-6 * h, tol = 1e-9 * h;
  struct key_t {
    long long a, b, c;
    bool operator==(const key_t& o) const { return a == o.a && b == o.b && c == o.c; }
  };
  struct hash_t {
    size_t operator()(const key_t& k) const {
      return std::hash<long long>()(k.a * 73856093LL ^ k.b * 19349663LL ^ k.c * 83492791LL);
    }
  };
  std::unordered_map<key_t, std::vector<size_t>, hash_t> grid;
  std::vector<size_t> rep(NL);
  for (size_t n = 0; n < NL; ++n) {
    const key_t k{(long long)std::floor(x[n] / cell), (long long)std::floor(y[n] / cell),
                  (long long)std::floor(z[n] / cell)};
    size_t found = n;
    for (int da = -1; da <= 1 && found == n; ++da)
      for (int db = -1; db <= 1 && found == n; ++db)
        for (int dc = -1; dc <= 1 && found == n; ++dc) {
          auto it = grid.find(key_t{k.a + da, k.b + db, k.c + dc});
          if (it == grid.end()) continue;
          for (size_t m : it->second)
            if (std::fabs(x[m] - x[n]) <= tol && std::fabs(y[m] - y[n]) <= tol && std::fabs(z[m] - z[n]) <= tol) {
              found = m;
              break;
            }
        }
    rep[n] = found;  // the first copy met (only first copies enter the grid): the smallest local index
    if (found == n) grid[k].push_back(n);
  }
  globalIds.malloc(NL);
  for (size_t n = 0; n < NL; ++n) globalIds[n] = 1 + (hlong)rep[n];
  // boundary faces by vertex set; face f of the reference hex (faces t=-1, s=-1, r=+1, s=+1, r=-1, t=+1)
  static const int fv[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 3, 7, 4}, {4, 5, 6, 7}};
  std::map<std::array<hlong, 4>, int> tagged;
  for (size_t b = 0; b + 4 < boundaryInfo.size(); b += 5) {
    std::array<hlong, 4> key{boundaryInfo[b + 1], boundaryInfo[b + 2], boundaryInfo[b + 3], boundaryInfo[b + 4]};
    std::sort(key.begin(), key.end());
    tagged[key] = (int)boundaryInfo[b];
  }
  std::vector<int> flag(NL, -1);
  auto onFace = [&](int f, int i, int j, int k) {
    switch (f) {
      case 0: return k == 0;
      case 1: return j == 0;
      case 2: return i == N;
      case 3: return j == N;
      case 4: return i == 0;
      default: return k == N;
    }
  };
  for (dlong e = 0; e < Nelements; ++e)
    for (int f = 0; f < 6; ++f) {
      std::array<hlong, 4> key;
      for (int v = 0; v < 4; ++v) key[v] = EToV[(size_t)e * 8 + fv[f][v]];
      std::sort(key.begin(), key.end());
      auto it = tagged.find(key);
      if (it == tagged.end() || it->second <= 0) continue;
      for (int k = 0; k < Nq; ++k)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i)
            if (onFace(f, i, j, k)) {
              const size_t r = rep[(size_t)e * Np + i + Nq * (j + Nq * k)];
              if (flag[r] < 0 || it->second < flag[r]) flag[r] = it->second;
            }
    }
  mapB.malloc(NL);
  for (size_t n = 0; n < NL; ++n) mapB[n] = flag[rep[n]];
}

void mesh_t::SetupBox() {
  hlong GX, GY, GZ;
  settings.getSetting("BOX GLOBAL NX", GX);
  settings.getSetting("BOX GLOBAL NY", GY);
  settings.getSetting("BOX GLOBAL NZ", GZ);
  settings.getSetting("BOX NX", nx);
  settings.getSetting("BOX NY", ny);
  settings.getSetting("BOX NZ", nz);
  if (GX * GY * GZ > 0) {  // global sizes override local ones (meshSetupBoxHex3D.cpp:59-65)
    nx = (dlong)GX;
    ny = (dlong)GY;
    nz = (dlong)GZ;
  }
  LIBP_ABORT("BOX NX/NY/NZ must be positive", nx <= 0 || ny <= 0 || nz <= 0);
  settings.getSetting("BOX BOUNDARY FLAG", boundaryFlag);
  periodic = (boundaryFlag == -1);
  // With two elements in a periodic direction both elements carry the same
  // eight vertex ids, so face matching by vertex set (meshConnect.cpp:85-110)
  // cannot tell their faces apart; the reference then merges unrelated nodes.
  LIBP_ABORT("periodic boxes need at least 3 elements per direction", periodic && (nx < 3 || ny < 3 || nz < 3));
  LIBP_ABORT("BOX BOUNDARY FLAG must be -1, 0, 1 or 2", boundaryFlag < -1 || boundaryFlag > 2);
  double DIMX, DIMY, DIMZ;
  settings.getSetting("BOX DIMX", DIMX);
  settings.getSetting("BOX DIMY", DIMY);
  settings.getSetting("BOX DIMZ", DIMZ);
  settings.getSetting("BOX CURVED MAP AMPLITUDE", curveAmp);
  const double dx = DIMX / nx, dy = DIMY / ny, dz = DIMZ / nz;
  const double X0 = -DIMX / 2.0, Y0 = -DIMY / 2.0, Z0 = -DIMZ / 2.0;
  const hlong NnX = periodic ? nx : nx + 1, NnY = periodic ? ny : ny + 1, NnZ = periodic ? nz : nz + 1;
  Nelements = nx * ny * nz;
  EToV.malloc((size_t)Nelements * Nverts);
  EX.malloc((size_t)Nelements * Nverts);
  EY.malloc((size_t)Nelements * Nverts);
  EZ.malloc((size_t)Nelements * Nverts);
  // vertex order of the reference hex (meshSetupBoxHex3D.cpp:126-152)
  static const int vi[8] = {0, 1, 1, 0, 0, 1, 1, 0}, vj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, vk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
#pragma omp parallel for collapse(3)
  for (dlong k = 0; k < nz; ++k)
    for (dlong j = 0; j < ny; ++j)
      for (dlong i = 0; i < nx; ++i) {
        const size_t e = i + (size_t)j * nx + (size_t)k * nx * ny;
        for (int v = 0; v < 8; ++v) {
          const hlong ii = (i + vi[v]) % NnX, jj = (j + vj[v]) % NnY, kk = (k + vk[v]) % NnZ;
          EToV[e * 8 + v] = ii + jj * NnX + kk * NnX * NnY;
          EX[e * 8 + v] = X0 + dx * i + (vi[v] ? dx : 0.0);
          EY[e * 8 + v] = Y0 + dy * j + (vj[v] ? dy : 0.0);
          EZ[e * 8 + v] = Z0 + dz * k + (vk[v] ? dz : 0.0);
        }
      }
  o_EX = platform.malloc(EX);
  o_EY = platform.malloc(EY);
  o_EZ = platform.malloc(EZ);
}

void mesh_t::ReferenceNodes(int NN) {
  LIBP_ABORT("POLYNOMIAL DEGREE must be in [1,15]", NN < 1 || NN > 15);
  N = NN;
  Nq = N + 1;
  Nfp = Nq * Nq;
  Np = Nq * Nq * Nq;
  gllNodesWeights(N, gllz, gllw);
  gllDmatrix(N, gllz, D);
  o_gllz = platform.malloc(gllz);
  o_gllw = platform.malloc(gllw);
  o_D = platform.malloc(D);
}

/* Node numbering. The reference labels every local node 1+(local index) and
 * propagates the minimum across face traces to a fixed point
 * (meshConnectNodes.cpp:32-117); on a structured box the fixed point is known:
 * a node's id is 1 + the smallest local index among its copies. Copies of a
 * lattice point lie in the elements of a product set of per-direction
 * candidates, so the smallest copy takes, in each direction, the lowest
 * element containing the point. mapB: the smallest positive boundary flag of
 * the boundary faces touching the node, else -1 (meshConnectNodes.cpp:52-67);
 * every box face carries the same flag. */
void mesh_t::ConnectNodes() {
  const size_t NL = (size_t)Nelements * Np;
  globalIds.malloc(NL);
  mapB.malloc(NL);
  const hlong LX = (hlong)nx * N, LY = (hlong)ny * N, LZ = (hlong)nz * N;
  auto lowest = [&](hlong X, hlong L, dlong nE, dlong& em, int& im) {
    if (periodic) X %= L;
    if (X == 0) {
      em = 0;
    } else if (X % N == 0) {
      em = (dlong)(X / N - 1);
    } else {
      em = (dlong)(X / N);
    }
    (void)nE;
    im = (int)(X - (hlong)em * N);
  };
  const bool flagged = !periodic && boundaryFlag > 0;
#pragma omp parallel for collapse(3)
  for (dlong ez = 0; ez < nz; ++ez)
    for (dlong ey = 0; ey < ny; ++ey)
      for (dlong ex = 0; ex < nx; ++ex) {
        const size_t e = ex + (size_t)ey * nx + (size_t)ez * nx * ny;
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i) {
              const hlong X = (hlong)ex * N + i, Y = (hlong)ey * N + j, Z = (hlong)ez * N + k;
              dlong exm, eym, ezm;
              int im, jm, km;
              lowest(X, LX, nx, exm, im);
              lowest(Y, LY, ny, eym, jm);
              lowest(Z, LZ, nz, ezm, km);
              const hlong em = exm + (hlong)eym * nx + (hlong)ezm * nx * ny;
              const size_t id = e * Np + i + j * Nq + k * Nq * Nq;
              globalIds[id] = 1 + em * Np + im + jm * Nq + km * Nq * Nq;
              const bool onB = (X == 0 || X == LX || Y == 0 || Y == LY || Z == 0 || Z == LZ);
              mapB[id] = (flagged && onB) ? boundaryFlag : -1;
            }
      }
}

void mesh_t::GeometricFactors() {
  const size_t NL = (size_t)Nelements * Np;
  o_x = platform.malloc<double>(NL);
  o_y = platform.malloc<double>(NL);
  o_z = platform.malloc<double>(NL);
  o_ggeo = platform.malloc<double>(NL * 6);
  o_wJ = platform.malloc<double>(NL);
  deviceMemory<int> o_bad = platform.malloc<int>(1);
  LIBP_HIP(hipMemsetAsync(o_bad.ptr(), 0, sizeof(int), platform.stream));
  LIBP_CALL(libp_geofacs_hex_f64(N, Nelements, o_EX.ptr(), o_EY.ptr(), o_EZ.ptr(), o_gllz.ptr(), o_gllw.ptr(),
                                 o_D.ptr(), curveAmp, o_x.ptr(), o_y.ptr(), o_z.ptr(), o_ggeo.ptr(), o_wJ.ptr(),
                                 o_bad.ptr(), platform.s()));
  int bad = 0;
  o_bad.copyTo(&bad, 1, 0, platform.stream);
  platform.finish();
  LIBP_ABORT("Negative J found at element " << bad - 1, bad > 0);
  o_ggeo32 = deviceMemory<float>();
  o_wJ32 = deviceMemory<float>();
}

void mesh_t::GatherScatterSetup() {
  NlocalGatherElements = Nelements;
  NglobalGatherElements = 0;
  localGatherElementList.malloc(Nelements);
  for (dlong e = 0; e < Nelements; ++e) localGatherElementList[e] = e;
}

void mesh_t::MakeFP32() {
  if (o_ggeo32.isInitialized()) return;
  const size_t NL = (size_t)Nelements * Np;
  o_ggeo32 = platform.malloc<float>(NL * 6);
  o_wJ32 = platform.malloc<float>(NL);
  o_D32 = platform.malloc<float>((size_t)Nq * Nq);
  LIBP_CALL(libp_convert_f64_to_f32((int)(NL * 6), o_ggeo.ptr(), o_ggeo32.ptr(), platform.s()));
  LIBP_CALL(libp_convert_f64_to_f32((int)NL, o_wJ.ptr(), o_wJ32.ptr(), platform.s()));
  LIBP_CALL(libp_convert_f64_to_f32(Nq * Nq, o_D.ptr(), o_D32.ptr(), platform.s()));
}

mesh_t mesh_t::SetupNewDegree(int Nf) {
  if (Nf == N) return *this;
  mesh_t m = *this;
  m.o_ggeo32 = deviceMemory<float>();
  m.o_wJ32 = deviceMemory<float>();
  m.o_D32 = deviceMemory<float>();
  m.ReferenceNodes(Nf);
  if (boxMesh) {
    m.ConnectNodes();
    m.GeometricFactors();
  } else {
    m.GeometricFactors();
    m.ConnectNodesGeneral();
  }
  m.GatherScatterSetup();
  return m;
}

}  // namespace libp
