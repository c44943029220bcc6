// elliptic_t for the C0 hex path on one MI355X GPU.
#include "elliptic.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace libp {

void elliptic_t::Setup(platform_t& _platform, mesh_t& _mesh, settings_t& _settings, dfloat _lambda,
                       const int _NBCTypes, const memory<int> _BCType, std::shared_ptr<randStream_t> _rng) {
  platform = _platform;
  mesh = _mesh;
  comm = _mesh.comm;
  settings = _settings;
  lambda = _lambda;
  rng = _rng ? _rng : std::make_shared<randStream_t>(1);
  disc_c0 = settings.compareSetting("DISCRETIZATION", "CONTINUOUS");
  LIBP_ABORT("Only DISCRETIZATION CONTINUOUS is supported on this path", !disc_c0);
  NBCTypes = _NBCTypes;
  BCType = _BCType.clone();
  BoundarySetup();
  SplitGather();
  SetupPrecon();
}

void elliptic_t::SetupPrecon() {
  if (settings.compareSetting("PRECONDITIONER", "JACOBI")) {
    precon.Setup<JacobiPrecon>(*this);
    // the fused Jacobi-PCG iteration's reduction partials, allocated here rather
    // than inside a timed solve
    if (settings.compareSetting("PCG FUSED OPERATOR", "TRUE") && mesh.N >= 1 && mesh.N <= 15) {
      o_axPartials = platform.malloc<double>(std::max(libp_ax_hex_split_blocks(mesh.N, mesh.NlocalGatherElements), 1));
      o_gatherPartials = platform.malloc<double>(std::max<long>(libp_pcg_gather_jacobi_partials(ogsMasked.NlocalT), 2));
    }
  } else if (settings.compareSetting("PRECONDITIONER", "MULTIGRID")) {
    precon.Setup<MultiGridPrecon>(*this);
  } else if (settings.compareSetting("PRECONDITIONER", "NONE")) {
    precon.Setup<IdentityPrecon>(Ndofs, platform);
  } else {
    LIBP_FORCE_ABORT("PRECONDITIONER " << settings.getSetting("PRECONDITIONER") << " is not supported on this path");
  }
}

void elliptic_t::BoundarySetup() {
  const dlong NL = mesh.Nelements * mesh.Np;
  split.reset();
  splitTried = false;
  allNeumann = (lambda == 0) ? 1 : 0;
  allNeumannPenalty = 1.;
  // node-wise flags (ellipticBoundarySetup.cpp:53-65); box faces all carry one flag,
  // so the face-wise EToB translation only matters through allNeumann
  if (mesh.boxMesh && !mesh.periodic && mesh.boundaryFlag > 0) {
    LIBP_ABORT("BOX BOUNDARY FLAG outside BCType table", mesh.boundaryFlag >= NBCTypes);
    if (BCType[mesh.boundaryFlag] != 2) allNeumann = 0;
  }
  if (!mesh.boxMesh) {  // any boundary node that is not Neumann (ellipticBoundarySetup.cpp:40-65)
    for (size_t n = 0; n < (size_t)NL; ++n) {
      const int bc = mesh.mapB[n];
      if (bc <= 0) continue;
      LIBP_ABORT("boundary flag " << bc << " outside BCType table", bc >= NBCTypes);
      if (BCType[bc] != 2) allNeumann = 0;
    }
  }
  Nmasked = 0;
  mapB.malloc(NL, 0);
  for (dlong n = 0; n < NL; ++n) {
    const int bc = mesh.mapB[n];
    if (bc > 0) {
      mapB[n] = BCType[bc];
      if (mapB[n] == 1) Nmasked++;
    }
  }
  maskIds.malloc(Nmasked);
  Nmasked = 0;
  for (dlong n = 0; n < NL; ++n)
    if (mapB[n] == 1) maskIds[Nmasked++] = n;
  o_mapB = platform.malloc(mapB);

  maskedGlobalIds = mesh.globalIds.clone();
  for (dlong n = 0; n < Nmasked; ++n) maskedGlobalIds[maskIds[n]] = 0;
  ogsMasked.Setup(NL, maskedGlobalIds, true, *rng, platform);
  allNeumannScale = 1. / std::sqrt((dfloat)ogsMasked.NgatherGlobal);

  // inverse multiplicity (ellipticBoundarySetup.cpp:102-114): the gather of
  // ones is exactly the row length
  weightG.malloc(ogsMasked.Ngather);
  for (dlong r = 0; r < ogsMasked.Ngather; ++r) {
    const double cnt = ogsMasked.rowStartsT[r + 1] - ogsMasked.rowStartsT[r];
    weightG[r] = (cnt > 0.0) ? 1. / cnt : 0.0;
  }
  o_weightG = platform.malloc(weightG);
  o_weightG32 = deviceMemory<float>();

  ogsMasked.SetupGlobalToLocalMapping(GlobalToLocal);
  o_GlobalToLocal = platform.malloc(GlobalToLocal);
  o_AqL = platform.malloc<dfloat>(NL);
  o_AqL32 = deviceMemory<float>();
  Ndofs = ogsMasked.Ngather * Nfields;
  Nhalo = 0;
}

void elliptic_t::PartialAx(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_AqL_) {
  if (mesh.N == 15) {  // 16x16 contractions on FP64 MFMA
    LIBP_CALL(libp_ax_hex_partial_mfma_n15(mesh.NlocalGatherElements, nullptr, o_GlobalToLocal.ptr(), mesh.o_wJ.ptr(),
                                           mesh.o_ggeo.ptr(), mesh.o_D.ptr(), lambda, o_q.ptr(), o_AqL_.ptr(),
                                           platform.s()));
    return;
  }
  LIBP_CALL(libp_ax_hex_partial_f64(mesh.N, mesh.NlocalGatherElements, nullptr, o_GlobalToLocal.ptr(),
                                    mesh.o_wJ.ptr(), mesh.o_ggeo.ptr(), mesh.o_D.ptr(), lambda, o_q.ptr(),
                                    o_AqL_.ptr(), platform.s()));
}

/* The split form of the operator: a row whose copy is interior to its element
 * has that one copy, so the Ax kernel stores it straight into Aq; the face
 * gather sums the rows of element-boundary nodes from the compact face
 * E-vector. Every row is one kind or the other on a conforming mesh; a mesh
 * where it is not aborts. Bit-identical to PartialAx + Gather. */
splitGather_t* elliptic_t::SplitGather() {
  if (splitTried) return split.get();
  splitTried = true;
  if (!settings.compareSetting("OPERATOR SPLIT GATHER", "TRUE") || mesh.N < 1 || mesh.N > 14) return nullptr;
  const int N = mesh.N, Nq = N + 1, Np = mesh.Np;
  const int Nface = libp_ax_hex_face_nodes(N);
  std::vector<int> faceIdx(Np, -1);
  for (int k = 0, f = 0; k < Nq; ++k)
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        const bool interior = i > 0 && i < N && j > 0 && j < N && k > 0 && k < N;
        if (!interior) faceIdx[i + j * Nq + k * Nq * Nq] = f++;
      }
  const memory<dlong>& rs = ogsMasked.rowStartsT;
  const memory<dlong>& ci = ogsMasked.colIdsT;
  const bool slotLayout = !settings.compareSetting("OPERATOR FACE LAYOUT", "ELEMENT");
  const size_t NfaceL = (size_t)mesh.Nelements * Nface;
  std::vector<dlong> rows, starts{0}, cols, slots;
  for (dlong r = 0; r < ogsMasked.Ngather; ++r) {
    const dlong a = rs[r], b = rs[r + 1];
    if (b - a == 1 && faceIdx[ci[a] % Np] < 0) continue;  // an interior node: written by the Ax kernel
    rows.push_back(r);
    for (dlong g = a; g < b; ++g) {
      const int f = faceIdx[ci[g] % Np];
      LIBP_ABORT("split operator: an element-interior node shares its row (non-conforming mesh)", f < 0);
      cols.push_back((ci[g] / Np) * Nface + f);
    }
    starts.push_back((dlong)cols.size());
  }
  const dlong Ncopies = (dlong)cols.size();
  if (slotLayout) {
    // each boundary node's slot in the row-ordered copy list; masked nodes (no row)
    // all write one trash slot past the end
    slots.assign(NfaceL, Ncopies);
    for (dlong s = 0; s < Ncopies; ++s) slots[cols[s]] = s;
  }
  auto sp = std::make_shared<splitGather_t>();
  sp->Nface = Nface;
  sp->NfaceRows = (dlong)rows.size();
  sp->slotLayout = slotLayout;
  sp->o_rows = platform.malloc<dlong>(std::max<size_t>(rows.size(), 1));
  if (!rows.empty()) sp->o_rows.copyFrom(rows.data(), rows.size(), 0, platform.stream);
  sp->o_rowStarts = platform.malloc<dlong>(starts.size());
  sp->o_rowStarts.copyFrom(starts.data(), starts.size(), 0, platform.stream);
  if (slotLayout) {
    sp->o_slots = platform.malloc<dlong>(std::max<size_t>(NfaceL, 1));
    if (NfaceL) sp->o_slots.copyFrom(slots.data(), NfaceL, 0, platform.stream);
    sp->o_AqFace = platform.malloc<dfloat>((size_t)Ncopies + 1);
  } else {
    sp->o_cols = platform.malloc<dlong>(std::max<size_t>(cols.size(), 1));
    if (!cols.empty()) sp->o_cols.copyFrom(cols.data(), cols.size(), 0, platform.stream);
    sp->o_AqFace = platform.malloc<dfloat>(NfaceL);
  }
  sp->nAxPartials = libp_ax_hex_split_blocks(N, mesh.NlocalGatherElements);
  sp->o_papPart = platform.malloc<double>(sp->nAxPartials + libp_ogs_gather_face_blocks(sp->NfaceRows) + 1);
  platform.finish();
  split = sp;
  return split.get();
}

void elliptic_t::Operator(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_Aq) {
  if (splitGather_t* sp = SplitGather()) {
    LIBP_CALL(libp_ax_hex_split_f64(mesh.N, mesh.NlocalGatherElements, nullptr, o_GlobalToLocal.ptr(), mesh.o_wJ.ptr(),
                                    mesh.o_ggeo.ptr(), mesh.o_D.ptr(), lambda, o_q.ptr(), o_Aq.ptr(),
                                    sp->o_AqFace.ptr(), sp->o_slots.ptr(), nullptr, platform.s()));
    LIBP_CALL(libp_ogs_gather_face_f64(sp->NfaceRows, sp->o_rows.ptr(), sp->o_rowStarts.ptr(), sp->o_cols.ptr(),
                                       sp->o_AqFace.ptr(), o_Aq.ptr(), platform.s()));
    return;
  }
  PartialAx(o_q, o_AqL);
  ogsMasked.Gather(o_Aq, o_AqL, 1, ogs::Add, ogs::Trans);
}

bool elliptic_t::OperatorJacobiPCG(deviceMemory<dfloat>& o_p, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r,
                                   deviceMemory<dfloat>& o_z, const deviceMemory<dfloat>& o_invDiag,
                                   double* scalars) {
  if (!settings.compareSetting("PCG FUSED OPERATOR", "TRUE") || mesh.N < 1 || mesh.N > 15 || SplitGather())
    return false;
  if (!o_gatherPartials.isInitialized()) {
    o_axPartials = platform.malloc<double>(std::max(libp_ax_hex_split_blocks(mesh.N, mesh.NlocalGatherElements), 1));
    o_gatherPartials = platform.malloc<double>(std::max<long>(libp_pcg_gather_jacobi_partials(ogsMasked.NlocalT), 2));
  }
  LIBP_CALL(libp_pcg_ax_hex_pap_f64(mesh.N, mesh.NlocalGatherElements, nullptr, o_GlobalToLocal.ptr(),
                                    mesh.o_wJ.ptr(), mesh.o_ggeo.ptr(), mesh.o_D.ptr(), lambda, o_p.ptr(),
                                    o_AqL.ptr(), o_axPartials.ptr(), scalars, platform.s()));
  LIBP_CALL(libp_pcg_gather_jacobi_update_f64(ogsMasked.NlocalT, ogsMasked.o_rowStartsT.ptr(),
                                              ogsMasked.o_colIdsT.ptr(), o_AqL.ptr(), o_invDiag.ptr(), o_p.ptr(),
                                              o_x.ptr(), o_r.ptr(), o_z.ptr(), scalars, o_gatherPartials.ptr(),
                                              platform.s()));
  return true;
}

bool elliptic_t::OperatorPAP(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_Aq, double* scalars, void* ws) {
  if (splitGather_t* sp = SplitGather()) {
    LIBP_CALL(libp_ax_hex_split_f64(mesh.N, mesh.NlocalGatherElements, nullptr, o_GlobalToLocal.ptr(), mesh.o_wJ.ptr(),
                                    mesh.o_ggeo.ptr(), mesh.o_D.ptr(), lambda, o_q.ptr(), o_Aq.ptr(),
                                    sp->o_AqFace.ptr(), sp->o_slots.ptr(), sp->o_papPart.ptr(), platform.s()));
    LIBP_CALL(libp_pcg_gather_face_pap_f64(sp->NfaceRows, sp->o_rows.ptr(), sp->o_rowStarts.ptr(), sp->o_cols.ptr(),
                                           sp->o_AqFace.ptr(), o_Aq.ptr(), o_q.ptr(), sp->nAxPartials,
                                           sp->o_papPart.ptr(), scalars, platform.s()));
    return true;
  }
  PartialAx(o_q, o_AqL);
  LIBP_CALL(libp_pcg_gather_pap_f64(ogsMasked.NlocalT, ogsMasked.o_rowStartsT.ptr(), ogsMasked.o_colIdsT.ptr(),
                                    o_AqL.ptr(), o_Aq.ptr(), o_q.ptr(), scalars, ws, platform.s()));
  return true;
}

double* elliptic_t::PartialAxLocal(deviceMemory<dfloat>& o_q) {
  PartialAx(o_q, o_AqL);
  return o_AqL.ptr();
}

float* elliptic_t::PartialAxLocalF32(deviceMemory<float>& o_q) {
  mesh.MakeFP32();
  if (!o_AqL32.isInitialized()) o_AqL32 = platform.malloc<float>((size_t)mesh.Nelements * mesh.Np);
  LIBP_CALL(libp_ax_hex_partial_f32(mesh.N, mesh.NlocalGatherElements, nullptr, o_GlobalToLocal.ptr(),
                                    mesh.o_wJ32.ptr(), mesh.o_ggeo32.ptr(), mesh.o_D32.ptr(), (float)lambda,
                                    o_q.ptr(), o_AqL32.ptr(), platform.s()));
  return o_AqL32.ptr();
}

void elliptic_t::OperatorResidual(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_res) {
  PartialAx(o_q, o_AqL);
  LIBP_CALL(libp_ogs_gather_residual_f64(ogsMasked.NlocalT, ogsMasked.o_rowStartsT.ptr(), ogsMasked.o_colIdsT.ptr(),
                                         o_AqL.ptr(), o_r.ptr(), o_res.ptr(), platform.s()));
}

void elliptic_t::OperatorResidualF32(deviceMemory<float>& o_q, deviceMemory<float>& o_r, deviceMemory<float>& o_res) {
  const float* AqL = PartialAxLocalF32(o_q);
  LIBP_CALL(libp_ogs_gather_residual_f32(ogsMasked.NlocalT, ogsMasked.o_rowStartsT.ptr(), ogsMasked.o_colIdsT.ptr(),
                                         AqL, o_r.ptr(), o_res.ptr(), platform.s()));
}

void elliptic_t::OperatorF32(deviceMemory<float>& o_q, deviceMemory<float>& o_Aq) {
  mesh.MakeFP32();
  if (!o_AqL32.isInitialized()) o_AqL32 = platform.malloc<float>((size_t)mesh.Nelements * mesh.Np);
  LIBP_CALL(libp_ax_hex_partial_f32(mesh.N, mesh.NlocalGatherElements, nullptr, o_GlobalToLocal.ptr(),
                                    mesh.o_wJ32.ptr(), mesh.o_ggeo32.ptr(), mesh.o_D32.ptr(), (float)lambda,
                                    o_q.ptr(), o_AqL32.ptr(), platform.s()));
  ogsMasked.Gather(o_Aq, o_AqL32, 1, ogs::Add, ogs::Trans);
}

void elliptic_t::BuildOperatorDiagonal(deviceMemory<dfloat>& o_diagA) {
  const dlong NL = mesh.Nelements * mesh.Np;
  deviceMemory<dfloat> o_diagL = platform.malloc<dfloat>(NL);
  const double shift = allNeumannPenalty * allNeumannScale * allNeumannScale;
  LIBP_CALL(libp_diag_hex_local_f64(mesh.N, mesh.Nelements, mesh.o_ggeo.ptr(), mesh.o_wJ.ptr(), mesh.o_D.ptr(),
                                    lambda, o_mapB.ptr(), allNeumann, shift, o_diagL.ptr(), platform.s()));
  ogsMasked.Gather(o_diagA, o_diagL, 1, ogs::Add, ogs::Trans);
  platform.finish();
}

void elliptic_t::ZeroMean(deviceMemory<dfloat>& o_q) {
  double* tmp = platform.o_scalars.ptr() + 8;
  LIBP_CALL(libp_sum_f64(Ndofs, o_q.ptr(), tmp, platform.o_reduceWS.ptr(), platform.s()));
  double qmean = platform.readScalars(tmp, 1)[0];
  qmean *= allNeumannScale * allNeumannScale;
  LIBP_CALL(libp_add_f64(Ndofs, -qmean, o_q.ptr(), platform.s()));
}

elliptic_t elliptic_t::SetupNewDegree(mesh_t& meshC) {
  if (meshC.N == mesh.N) return *this;
  elliptic_t e = *this;
  e.mesh = meshC;
  e.BoundarySetup();
  e.precon = precon_t();
  return e;
}

int elliptic_t::dataKind() const {
  const std::string f = settings.getSetting("DATA FILE");
  if (f.find("Fourier3D") != std::string::npos) return 1;
  LIBP_ABORT("DATA FILE " << f << " is not supported (ellipticSine3D.h or ellipticFourier3D)",
             f.find("Sine3D") == std::string::npos);
  return 0;
}

// 64-mode forcing sum_m a_m sin(pi k_m.x + phi_m): k in {1..4}^3, a = +-1/|k|^2,
// phi uniform in [0,2pi), drawn with erand48 from [DATA SEED].
deviceMemory<dfloat> elliptic_t::dataCoefficients() {
  long seed = 1;
  settings.getSetting("DATA SEED", seed);
  return dataCoefficients(seed);
}

deviceMemory<dfloat> elliptic_t::dataCoefficients(long seed) {
  if (dataKind() != 1) return deviceMemory<dfloat>();
  unsigned short xs[3] = {0x330E, (unsigned short)(seed & 0xffff), (unsigned short)((seed >> 16) & 0xffff)};
  std::vector<double> c(64 * 5);
  for (int m = 0; m < 64; ++m) {
    const int kx = 1 + (int)(4 * erand48(xs));
    const int ky = 1 + (int)(4 * erand48(xs));
    const int kz = 1 + (int)(4 * erand48(xs));
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double sg = (erand48(xs) < 0.5) ? -1.0 : 1.0;
    c[5 * m + 0] = kx;
    c[5 * m + 1] = ky;
    c[5 * m + 2] = kz;
    c[5 * m + 3] = sg / k2;
    c[5 * m + 4] = 2 * M_PI * erand48(xs);
  }
  return platform.malloc(c);
}

int elliptic_t::Solve(linearSolver_t& linearSolver, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r,
                      const dfloat tol, const int MAXIT, const int verbose) {
  if (allNeumann) ZeroMean(o_r);
  return linearSolver.Solve(*this, precon, o_x, o_r, tol, MAXIT, verbose);
}

void elliptic_t::BuildRhs(long seed, deviceMemory<dfloat>& o_r) {
  const int kind = dataKind();
  deviceMemory<dfloat> o_coeffs = dataCoefficients(seed);
  const dlong NL = mesh.Nelements * mesh.Np;
  deviceMemory<dfloat> o_rL = platform.malloc<dfloat>(NL), o_ubL = platform.malloc<dfloat>(NL);
  deviceMemory<dfloat> o_AubL = platform.malloc<dfloat>(NL);
  LIBP_CALL(libp_rhs_hex_f64(mesh.Nelements, mesh.Np, kind, o_coeffs.ptr(), mesh.o_x.ptr(), mesh.o_y.ptr(),
                             mesh.o_z.ptr(), mesh.o_wJ.ptr(), lambda, o_mapB.ptr(), o_rL.ptr(), o_ubL.ptr(),
                             platform.s()));
  LIBP_CALL(libp_ax_hex_local_f64(mesh.N, mesh.Nelements, mesh.o_wJ.ptr(), mesh.o_ggeo.ptr(), mesh.o_D.ptr(), lambda,
                                  o_ubL.ptr(), o_AubL.ptr(), platform.s()));
  LIBP_CALL(libp_axpy_f64(NL, -1.0, o_AubL.ptr(), 1.0, o_rL.ptr(), platform.s()));
  ogsMasked.Gather(o_r, o_rL, 1, ogs::Add, ogs::Trans);
  platform.finish();
}

ellipticRunResult_t elliptic_t::Run(int verboseOverride) {
  ellipticRunResult_t res;
  const hlong NglobalDofs = ogsMasked.NgatherGlobal * Nfields;
  // solver choice in the order of ellipticRun.cpp:41-51 (compareSetting is a substring test)
  linearSolver_t linearSolver;
  if (settings.compareSetting("LINEAR SOLVER", "NBPCG")) {
    linearSolver.Setup<LinearSolver::nbpcg>(Ndofs, Nhalo, platform, settings, comm);
  } else if (settings.compareSetting("LINEAR SOLVER", "NBFPCG")) {
    linearSolver.Setup<LinearSolver::nbfpcg>(Ndofs, Nhalo, platform, settings, comm);
  } else if (settings.compareSetting("LINEAR SOLVER", "PCG")) {
    linearSolver.Setup<LinearSolver::pcg>(Ndofs, Nhalo, platform, settings, comm);
  } else {
    LIBP_FORCE_ABORT("LINEAR SOLVER " << settings.getSetting("LINEAR SOLVER") << " is not supported on this path");
  }
  linearSolver.get()->pureRelative = settings.compareSetting("LINEAR SOLVER STOPPING CRITERION", "RELATIVE") &&
                                     !settings.compareSetting("LINEAR SOLVER STOPPING CRITERION", "ABS/REL");

  const int kind = dataKind();
  deviceMemory<dfloat> o_coeffs = dataCoefficients();
  const dlong NL = mesh.Nelements * mesh.Np;
  deviceMemory<dfloat> o_rL = platform.malloc<dfloat>(NL), o_xL = platform.malloc<dfloat>(NL);
  deviceMemory<dfloat> o_ubL = platform.malloc<dfloat>(NL), o_AubL = platform.malloc<dfloat>(NL);
  LIBP_CALL(libp_set_f64(NL, 0.0, o_xL.ptr(), platform.s()));
  // forcing (ellipticRhsHex3D.okl) and Dirichlet lift (ellipticRhsBCHex3D.okl: rhs -= A_L u_b)
  LIBP_CALL(libp_rhs_hex_f64(mesh.Nelements, mesh.Np, kind, o_coeffs.ptr(), mesh.o_x.ptr(), mesh.o_y.ptr(),
                             mesh.o_z.ptr(), mesh.o_wJ.ptr(), lambda, o_mapB.ptr(), o_rL.ptr(), o_ubL.ptr(),
                             platform.s()));
  LIBP_CALL(libp_ax_hex_local_f64(mesh.N, mesh.Nelements, mesh.o_wJ.ptr(), mesh.o_ggeo.ptr(), mesh.o_D.ptr(), lambda,
                                  o_ubL.ptr(), o_AubL.ptr(), platform.s()));
  LIBP_CALL(libp_axpy_f64(NL, -1.0, o_AubL.ptr(), 1.0, o_rL.ptr(), platform.s()));

  deviceMemory<dfloat> o_r = platform.malloc<dfloat>(Ndofs + Nhalo), o_x = platform.malloc<dfloat>(Ndofs + Nhalo);
  ogsMasked.Gather(o_r, o_rL, 1, ogs::Add, ogs::Trans);
  ogsMasked.Gather(o_x, o_xL, 1, ogs::Add, ogs::NoTrans);

  int maxIter = 5000;
  settings.getSetting("MAXIMUM ITERATIONS", maxIter);
  double tol = 1e-8;
  settings.getSetting("LINEAR SOLVER TOLERANCE", tol);
  const int verbose = (verboseOverride >= 0) ? verboseOverride : (settings.compareSetting("VERBOSE", "TRUE") ? 1 : 0);

  // x is the gather of a zero E-vector: the solver's initial residual is the rhs
  linearSolver.get()->zeroInitialGuess = true;
  platform.finish();
  const auto t0 = std::chrono::steady_clock::now();
  const int iter = Solve(linearSolver, o_x, o_r, tol, maxIter, verbose);
  ogsMasked.Scatter(o_xL, o_x, 1, ogs::NoTrans);
  LIBP_CALL(libp_addbc_hex_f64(NL, kind, o_coeffs.ptr(), mesh.o_x.ptr(), mesh.o_y.ptr(), mesh.o_z.ptr(),
                               o_mapB.ptr(), o_xL.ptr(), platform.s()));
  platform.finish();
  const auto t1 = std::chrono::steady_clock::now();
  const double elapsed = std::chrono::duration<double>(t1 - t0).count();
  if (verbose)
    printf("%d, %lld, %g, %d, %g, %g; global: N, dofs, elapsed, iterations, time per node, nodes*iterations/time %s\n",
           mesh.N, (long long)NglobalDofs, elapsed, iter, elapsed / NglobalDofs, NglobalDofs * ((dfloat)iter / elapsed),
           settings.getSetting("PRECONDITIONER").c_str());

  // Solution norm sqrt(xL' M xL) (ellipticRun.cpp:238-247)
  double* tmp = platform.o_scalars.ptr() + 8;
  LIBP_CALL(libp_weighted_inner_prod_f64(NL, mesh.o_wJ.ptr(), o_xL.ptr(), o_xL.ptr(), tmp, platform.o_reduceWS.ptr(),
                                         platform.s()));
  res.norm = std::sqrt(platform.readScalars(tmp, 1)[0]);
  if (kind == 0) {
    // L2 error against the manufactured solution: u_b evaluated at every node
    memory<int> all(NL, 1);
    deviceMemory<int> o_all = platform.malloc(all);
    LIBP_CALL(libp_addbc_hex_f64(NL, 0, nullptr, mesh.o_x.ptr(), mesh.o_y.ptr(), mesh.o_z.ptr(), o_all.ptr(),
                                 o_ubL.ptr(), platform.s()));
    LIBP_CALL(libp_axpy_f64(NL, 1.0, o_xL.ptr(), -1.0, o_ubL.ptr(), platform.s()));
    LIBP_CALL(libp_weighted_inner_prod_f64(NL, mesh.o_wJ.ptr(), o_ubL.ptr(), o_ubL.ptr(), tmp,
                                           platform.o_reduceWS.ptr(), platform.s()));
    res.maxError = std::sqrt(platform.readScalars(tmp, 1)[0]);
  }
  // field file (ellipticRun.cpp:223-235)
  if (settings.compareSetting("OUTPUT TO FILE", "TRUE")) {
    std::string name = settings.getSetting("OUTPUT FILE NAME");
    char fname[4096];
    std::snprintf(fname, sizeof(fname), "%s_%04d.vtu", name.c_str(), comm.rank());
    PlotFields(o_xL, fname);
  }
  if (comm.rank() == 0 && verbose) printf("Solution norm = %17.15lg\n", res.norm);
  res.iterations = iter;
  res.solveSeconds = elapsed;
  res.r0 = std::sqrt(linearSolver.get()->stats.rdotr0);
  res.rFinal = std::sqrt(std::max(linearSolver.get()->stats.rdotrFinal, 0.0));
  o_xL_last = o_xL;
  return res;
}

/* Plot output in the reference's VTU layout (ellipticPlotFields.cpp,
 * meshReferenceNodesHex3D.cpp:54-67, meshBasisHex3D.cpp EquispacedEToVHex3D):
 * each element's nodal values are interpolated (Lagrange, tensor product) to
 * plotNq = N+4 equispaced nodes per direction; every sub-cube is split into
 * the same six tetrahedra (VTK type 10) with corner c = i + 2j + 4k:
 * (0,3,2,7) (0,1,3,7) (0,2,6,7) (0,6,4,7) (0,5,1,7) (0,4,5,7). */
void elliptic_t::PlotFields(const deviceMemory<dfloat>& o_qL, const std::string& fileName) {
  const int N = mesh.N, Nq = mesh.Nq, Np = mesh.Np;
  const int pN = N + 3, pNq = pN + 1, pNp = pNq * pNq * pNq;
  const dlong E = mesh.Nelements;
  const size_t NL = (size_t)E * Np;
  std::vector<double> x(NL), y(NL), z(NL), q(NL);
  mesh.o_x.copyTo(x.data(), NL, 0, platform.stream);
  mesh.o_y.copyTo(y.data(), NL, 0, platform.stream);
  mesh.o_z.copyTo(z.data(), NL, 0, platform.stream);
  o_qL.copyTo(q.data(), NL, 0, platform.stream);
  platform.finish();
  // I[i][m] = l_m(r_i): Lagrange basis of the GLL nodes at the equispaced nodes
  std::vector<double> I((size_t)pNq * Nq);
  for (int i = 0; i < pNq; ++i) {
    const double r = -1.0 + 2.0 * i / pN;
    for (int m = 0; m < Nq; ++m) {
      double l = 1.0;
      for (int k = 0; k < Nq; ++k)
        if (k != m) l *= (r - mesh.gllz[k]) / (mesh.gllz[m] - mesh.gllz[k]);
      I[(size_t)i * Nq + m] = l;
    }
  }
  std::vector<double> t1((size_t)pNq * Nq * Nq), t2((size_t)pNq * pNq * Nq);
  auto interp = [&](const double* u, double* out) {
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < pNq; ++i) {
          double s = 0;
          for (int m = 0; m < Nq; ++m) s += I[(size_t)i * Nq + m] * u[m + Nq * (j + Nq * k)];
          t1[i + pNq * (j + Nq * k)] = s;
        }
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < pNq; ++j)
        for (int i = 0; i < pNq; ++i) {
          double s = 0;
          for (int m = 0; m < Nq; ++m) s += I[(size_t)j * Nq + m] * t1[i + pNq * (m + Nq * k)];
          t2[i + pNq * (j + pNq * k)] = s;
        }
    for (int k = 0; k < pNq; ++k)
      for (int j = 0; j < pNq; ++j)
        for (int i = 0; i < pNq; ++i) {
          double s = 0;
          for (int m = 0; m < Nq; ++m) s += I[(size_t)k * Nq + m] * t2[i + pNq * (j + pNq * m)];
          out[i + pNq * (j + pNq * k)] = s;
        }
  };
  FILE* fp = std::fopen(fileName.c_str(), "w");
  LIBP_ABORT("cannot open " << fileName, !fp);
  const long Ncells = (long)E * 6 * pN * pN * pN;
  std::fprintf(fp, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"BigEndian\">\n");
  std::fprintf(fp, "  <UnstructuredGrid>\n");
  std::fprintf(fp, "    <Piece NumberOfPoints=\"%ld\" NumberOfCells=\"%ld\">\n", (long)E * pNp, Ncells);
  std::fprintf(fp, "      <Points>\n");
  std::fprintf(fp, "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n");
  std::vector<double> px(pNp), py(pNp), pz(pNp), pq(pNp);
  for (dlong e = 0; e < E; ++e) {
    interp(&x[(size_t)e * Np], px.data());
    interp(&y[(size_t)e * Np], py.data());
    interp(&z[(size_t)e * Np], pz.data());
    for (int n = 0; n < pNp; ++n) std::fprintf(fp, "       %g %g %g\n", px[n], py[n], pz[n]);
  }
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "      </Points>\n");
  std::fprintf(fp, "      <PointData Scalars=\"scalars\">\n");
  std::fprintf(fp, "        <DataArray type=\"Float32\" Name=\"Fields\" NumberOfComponents=\"%d\" Format=\"ascii\">\n",
               Nfields);
  for (dlong e = 0; e < E; ++e) {
    interp(&q[(size_t)e * Np], pq.data());
    for (int n = 0; n < pNp; ++n) std::fprintf(fp, "       %g \n", pq[n]);
  }
  std::fprintf(fp, "       </DataArray>\n");
  std::fprintf(fp, "     </PointData>\n");
  std::fprintf(fp, "    <Cells>\n");
  std::fprintf(fp, "      <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n");
  static const int tets[6][4] = {{0, 3, 2, 7}, {0, 1, 3, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 5, 1, 7}, {0, 4, 5, 7}};
  for (dlong e = 0; e < E; ++e)
    for (int k = 0; k < pN; ++k)
      for (int j = 0; j < pN; ++j)
        for (int i = 0; i < pN; ++i)
          for (int t = 0; t < 6; ++t) {
            std::fprintf(fp, "       ");
            for (int v = 0; v < 4; ++v) {
              const int c = tets[t][v];
              const int n = (i + (c & 1)) + pNq * ((j + ((c >> 1) & 1)) + pNq * (k + ((c >> 2) & 1)));
              std::fprintf(fp, "%ld ", (long)e * pNp + n);
            }
            std::fprintf(fp, "\n");
          }
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n");
  for (long c = 1; c <= Ncells; ++c) std::fprintf(fp, "       %ld\n", 4 * c);
  std::fprintf(fp, "       </DataArray>\n");
  std::fprintf(fp, "       <DataArray type=\"Int32\" Name=\"types\" Format=\"ascii\">\n");
  for (long c = 0; c < Ncells; ++c) std::fprintf(fp, "10\n");
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "      </Cells>\n");
  std::fprintf(fp, "    </Piece>\n");
  std::fprintf(fp, "  </UnstructuredGrid>\n");
  std::fprintf(fp, "</VTKFile>\n");
  std::fclose(fp);
}

/* JacobiPrecon (ellipticPreconJacobi.cpp:30-51). */
JacobiPrecon::JacobiPrecon(elliptic_t& _elliptic) : elliptic(_elliptic) {
  deviceMemory<dfloat> o_diagA = elliptic.platform.malloc<dfloat>(elliptic.Ndofs);
  elliptic.BuildOperatorDiagonal(o_diagA);
  std::vector<double> d(elliptic.Ndofs);
  o_diagA.copyTo(d.data(), d.size(), 0, elliptic.platform.stream);
  elliptic.platform.finish();
  for (auto& v : d) v = 1.0 / v;
  o_invDiagA = elliptic.platform.malloc(d);
}

void JacobiPrecon::Operator(deviceMemory<dfloat>& o_r, deviceMemory<dfloat>& o_Mr) {
  LIBP_CALL(libp_amxpy_f64(elliptic.Ndofs, 1.0, o_invDiagA.ptr(), o_r.ptr(), 0.0, o_Mr.ptr(), elliptic.platform.s()));
  if (elliptic.allNeumann) elliptic.ZeroMean(o_Mr);
}

std::vector<int> multigridDegrees(const settings_t& settings, int N) {
  std::vector<int> deg;
  if (settings.compareSetting("MULTIGRID COARSENING", "CUSTOM")) {
    std::string v = settings.getSetting("MULTIGRID DEGREES");
    size_t p = 0;
    while (p < v.size()) {
      size_t q = v.find(',', p);
      if (q == std::string::npos) q = v.size();
      deg.push_back(std::atoi(v.substr(p, q - p).c_str()));
      p = q + 1;
    }
    LIBP_ABORT("MULTIGRID DEGREES must start at the POLYNOMIAL DEGREE and end at 1",
               deg.empty() || deg.front() != N || deg.back() != 1);
    for (size_t i = 1; i < deg.size(); ++i) LIBP_ABORT("MULTIGRID DEGREES must decrease", deg[i] >= deg[i - 1]);
    return deg;
  }
  int Nf = N, Nc = N, NpFine = (N + 1) * (N + 1) * (N + 1), NpCoarse = NpFine;
  deg.push_back(N);
  while (Nc > 1) {
    if (settings.compareSetting("MULTIGRID COARSENING", "ALLDEGREES")) {
      Nc = Nf - 1;
    } else if (settings.compareSetting("MULTIGRID COARSENING", "HALFDEGREES")) {
      Nc = std::max(1, (Nf + 1) / 2);
    } else {
      while (NpCoarse > NpFine / 2 && Nc > 1) {
        Nc--;
        NpCoarse = (Nc + 1) * (Nc + 1) * (Nc + 1);
      }
    }
    NpCoarse = (Nc + 1) * (Nc + 1) * (Nc + 1);
    deg.push_back(Nc);
    Nf = Nc;
    NpFine = NpCoarse;
  }
  return deg;
}

}  // namespace libp
