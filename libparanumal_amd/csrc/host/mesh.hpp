// Hex box mesh for the MI355X elliptic path (mesh_t subset).
#pragma once

#include <string>
#include <vector>

#include "core.hpp"

namespace libp {

/* 1-D GLL basis by closed forms: nodes are the roots of (1-x^2) P'_N found by
 * Newton iteration, weights 2/(N(N+1) P_N(x)^2), derivative matrix
 * D[i*Nq+m] = l'_m(r_i) from the GLL closed form, and the Lagrange degree-raise
 * matrix P[f*NqC+c] = l^c_c(r_f). They stand in for the Vandermonde/LAPACK
 * route of libs/mesh/meshBasis1D.cpp:114-372 (equal to rounding). */
void gllNodesWeights(int N, std::vector<double>& r, std::vector<double>& w);
void gllDmatrix(int N, const std::vector<double>& r, std::vector<double>& D);
void degreeRaiseMatrix(int Nc, int Nf, std::vector<double>& P);

class mesh_t {
 public:
  platform_t platform;
  settings_t settings;
  comm_t comm;

  static constexpr int Nverts = 8, Nfaces = 6, NfaceVertices = 4;
  int N = 0, Nq = 0, Np = 0, Nfp = 0;
  dlong Nelements = 0;
  dlong nx = 0, ny = 0, nz = 0;
  int boundaryFlag = 1;
  bool periodic = false;
  bool boxMesh = true;  // BOX (closed-form numbering) or a mesh file (general numbering)
  // mesh files: boundary faces as (physical tag, 4 vertex ids) (meshReadGmshHex3D.cpp)
  std::vector<hlong> boundaryInfo;
  double curveAmp = 0.0;

  memory<double> EX, EY, EZ;  // [e][8]
  memory<hlong> EToV;         // [e][8]
  std::vector<double> gllz, gllw, D;

  memory<hlong> globalIds;  // host, [Nelements*Np]
  memory<int> mapB;         // host, [Nelements*Np]

  dlong NlocalGatherElements = 0, NglobalGatherElements = 0;
  memory<dlong> localGatherElementList;

  deviceMemory<double> o_EX, o_EY, o_EZ, o_gllz, o_gllw, o_D;
  deviceMemory<double> o_x, o_y, o_z, o_ggeo, o_wJ;
  deviceMemory<float> o_D32, o_ggeo32, o_wJ32;

  mesh_t() {}
  mesh_t(platform_t& platform, settings_t& settings, comm_t comm) { Setup(platform, settings, comm); }

  // meshSetup.cpp:31-102 (BOX meshes and Gmsh hexahedral mesh files)
  void Setup(platform_t& platform, settings_t& settings, comm_t comm);
  // meshSetupNewDegree.cpp:32-64
  mesh_t SetupNewDegree(int Nf);

  void SetupBox();           // meshSetupBoxHex3D.cpp:31-232
  void ReadGmshHex3D(const std::string& fileName);  // meshReadGmshHex3D.cpp
  void ConnectNodesGeneral();  // numbering and boundary flags of a mesh file (see mesh.cpp)
  void ReferenceNodes(int N);
  void ConnectNodes();       // globalIds + mapB in closed form (see mesh.cpp)
  void GeometricFactors();   // physical nodes + ggeo + wJ on the device
  void GatherScatterSetup(); // single rank: all elements are local gather elements
  void MakeFP32();           // single-precision copies for the p-MG smoother
};

}  // namespace libp
