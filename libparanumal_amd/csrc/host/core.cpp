// Core runtime implementation: settings and the HIP platform.
#include "core.hpp"

#include <cctype>
#include <iostream>

namespace libp {

void settings_t::newSetting(const std::string& name, const std::string& val, const std::string& description,
                            const std::vector<std::string>& options) {
  LIBP_ABORT("Setting with name: [" << name << "] already exists.", settings.count(name));
  settings[name] = setting_t{name, val, description, options};
  insertOrder.push_back(name);
}

// setting_t::updateVal (settings.cpp:41-59): options, when given, are exhaustive.
void settings_t::changeSetting(const std::string& name, const std::string& newVal) {
  auto it = settings.find(name);
  LIBP_ABORT("Setting with name: [" << name << "] does not exist.", it == settings.end());
  setting_t& s = it->second;
  if (!s.options.empty()) {
    bool ok = false;
    for (auto& o : s.options) ok = ok || (o == newVal);
    LIBP_ABORT("Value: \"" << newVal << "\" not valid for setting " << name, !ok);
  }
  s.val = newVal;
}

std::string settings_t::getSetting(const std::string& name) const {
  auto it = settings.find(name);
  LIBP_ABORT("Unable to find setting: [" << name << "]", it == settings.end());
  return it->second.val;
}

bool settings_t::compareSetting(const std::string& name, const std::string& token) const {
  return getSetting(name).find(token) != std::string::npos;
}

// Line grammar of settings_t::readSettingsFromFile (settings.cpp:134-214):
// '#' ends the line, "[NAME]" starts a setting, all other non-space characters
// accumulate into the value; a pair is emitted when both parts are non-empty.
std::vector<std::pair<std::string, std::string>> settings_t::parseText(const std::string& text) {
  std::vector<std::pair<std::string, std::string>> out;
  std::istringstream in(text);
  std::string line, name, val;
  while (std::getline(in, line)) {
    LIBP_ABORT("Line in settings file is too long: " << line, line.size() > 512);
    for (size_t i = 0; i < line.size(); ++i) {
      const char c = line[i];
      if (c == '#') break;
      if (c == '[') {
        if (!name.empty() && !val.empty()) {
          out.emplace_back(name, val);
          name.clear();
          val.clear();
        }
        name.clear();
        val.clear();
        ++i;
        while (i < line.size() && line[i] != ']') name += line[i++];
      } else if (!isspace((unsigned char)c)) {
        val += c;
      }
    }
    if (!name.empty() && !val.empty()) {
      out.emplace_back(name, val);
      name.clear();
      val.clear();
    }
  }
  return out;
}

// ellipticSettings_t::parseFromFile (ellipticSettings.cpp:127-148): every key
// must already exist.
void settings_t::readSettingsFromText(const std::string& text) {
  for (auto& kv : parseText(text)) {
    LIBP_ABORT("Unknown setting: [" << kv.first << "] requested", !hasSetting(kv.first));
    changeSetting(kv.first, kv.second);
  }
}

void settings_t::report(std::ostream& os) const {
  for (auto& n : insertOrder) os << "[" << n << "] " << settings.at(n).val << "\n";
}

settings_t ellipticDefaultSettings() {
  settings_t s;
  // platformSettings_t (libs/core/platformSettings.cpp:31-56)
  s.newSetting("FORMAT", "2.0", "Setting file version number", {"2.0"});
  s.newSetting("THREAD MODEL", "HIP", "Parallel execution platform", {"Serial", "OpenMP", "CUDA", "HIP", "OpenCL"});
  s.newSetting("PLATFORM NUMBER", "0");
  s.newSetting("DEVICE NUMBER", "0");
  s.newSetting("CACHE DIR", ".");
  // meshSettings_t (libs/mesh/meshSettings.cpp:32-91)
  s.newSetting("MESH FILE", "BOX");
  s.newSetting("MESH DIMENSION", "2", "Dimension of input mesh", {"2", "3"});
  s.newSetting("ELEMENT TYPE", "3", "Type of mesh elements (number of edges)", {"3", "4", "6", "12"});
  s.newSetting("ELEMENT MAP", "ISOPARAMETRIC", "Element map", {"ISOPARAMETRIC", "AFFINE"});
  s.newSetting("BOX DIMX", "10");
  s.newSetting("BOX DIMY", "10");
  s.newSetting("BOX DIMZ", "10");
  s.newSetting("BOX NX", "10");
  s.newSetting("BOX NY", "10");
  s.newSetting("BOX NZ", "10");
  s.newSetting("BOX GLOBAL NX", "0");
  s.newSetting("BOX GLOBAL NY", "0");
  s.newSetting("BOX GLOBAL NZ", "0");
  s.newSetting("BOX BOUNDARY FLAG", "1");
  s.newSetting("POLYNOMIAL DEGREE", "4", "Degree of polynomial finite element space",
               {"1", "2", "3", "4", "5", "6", "7", "8", "9", "10", "11", "12", "13", "14", "15"});
  // paradogs settings accepted for file compatibility (meshSettings.cpp:90)
  s.newSetting("PARADOGS PARTITIONING", "INERTIAL");
  // this build: smooth node perturbation x += A sin(2 pi y) sin(2 pi z)
  s.newSetting("BOX CURVED MAP AMPLITUDE", "0", "Amplitude of the curved box map (0 = trilinear box)");
  // ellipticAddRunSettings (ellipticSettings.cpp:38-54)
  s.newSetting("DATA FILE", "data/ellipticSine2D.h", "Boundary and Initial conditions header");
  s.newSetting("LAMBDA", "1.0");
  s.newSetting("OUTPUT TO FILE", "FALSE", "", {"TRUE", "FALSE"});
  s.newSetting("OUTPUT FILE NAME", "elliptic");
  // ellipticAddSettings (ellipticSettings.cpp:56-103)
  s.newSetting("DISCRETIZATION", "CONTINUOUS", "", {"CONTINUOUS", "IPDG"});
  s.newSetting("LINEAR SOLVER", "PCG", "", {"PCG", "FPCG", "NBPCG", "NBFPCG", "PGMRES", "PMINRES"});
  s.newSetting("LINEAR SOLVER STOPPING CRITERION", "ABS/REL-INITRESID", "",
               {"ABS/REL-INITRESID", "ABS/REL-RHS-2NORM", "RELATIVE"});
  s.newSetting("PRECONDITIONER", "NONE", "",
               {"NONE", "JACOBI", "MASSMATRIX", "PARALMOND", "MULTIGRID", "SEMFEM", "OAS"});
  s.newSetting("MULTIGRID COARSENING", "HALFDOFS", "", {"ALLDEGREES", "HALFDEGREES", "HALFDOFS", "CUSTOM"});
  s.newSetting("MULTIGRID SMOOTHER", "CHEBYSHEV", "", {"DAMPEDJACOBI", "CHEBYSHEV", "LEARNED"});
  s.newSetting("MULTIGRID CHEBYSHEV DEGREE", "2");
  s.newSetting("VERBOSE", "FALSE", "", {"TRUE", "FALSE"});
  // parAlmond (libs/parAlmond/parAlmondSettings.cpp:33-60)
  s.newSetting("PARALMOND CYCLE", "VCYCLE", "", {"VCYCLE", "KCYCLE", "EXACT"});
  s.newSetting("PARALMOND STRENGTH", "SYMMETRIC", "", {"RUGESTUBEN", "SYMMETRIC"});
  s.newSetting("PARALMOND AGGREGATION", "SMOOTHED", "", {"SMOOTHED", "UNSMOOTHED"});
  s.newSetting("PARALMOND SMOOTHER", "CHEBYSHEV", "", {"DAMPEDJACOBI", "CHEBYSHEV"});
  s.newSetting("PARALMOND CHEBYSHEV DEGREE", "2");
  // this build: p-MG degree list, precision, coarse Jacobi-PCG and solver knobs
  s.newSetting("MULTIGRID DEGREES", "", "Comma-separated degrees for MULTIGRID COARSENING CUSTOM, e.g. 7,3,1");
  s.newSetting("MULTIGRID PRECISION", "FP32", "Precision of p-MG smoothing and transfers", {"FP32", "FP64"});
  s.newSetting("MULTIGRID COARSE TOLERANCE", "1e-2", "Relative tolerance of the coarse Jacobi-PCG");
  s.newSetting("MULTIGRID COARSE MAXIT", "200");
  s.newSetting("MULTIGRID COARSE FUSED", "TRUE", "Coarse Jacobi-PCG at N=1 in three fused kernels per iteration",
               {"TRUE", "FALSE"});
  s.newSetting("MULTIGRID COARSE GRAPH", "TRUE", "Replay the fused coarse iterations as captured HIP graphs",
               {"TRUE", "FALSE"});
  s.newSetting("PCG JACOBI FUSED", "TRUE",
               "(F)PCG in fused launches: p.Ap in the operator's gather; with Jacobi, z = invDiag.*r and r.z in "
               "the update",
               {"TRUE", "FALSE"});
  s.newSetting("LINEAR SOLVER CHECK EVERY", "4",
               "(F)PCG with a Jacobi preconditioner: queue this many iterations between host reads of r.r; the "
               "kernels stop themselves at the tolerance (device STOP flag), so the iterates and the iteration "
               "count are those of a test after every iteration");
  s.newSetting("PCG FUSED OPERATOR", "TRUE",
               "PCG + Jacobi in three passes per iteration: the p update, the Ax kernel with p.Ap, and the gather "
               "fused with the x, r, z update (N <= 15, two-pass operator); FALSE runs the gather (with p.Ap) and "
               "the update as separate passes",
               {"TRUE", "FALSE"});
  s.newSetting("OPERATOR SPLIT GATHER", "FALSE",
               "FP64 operator with element-interior rows written by the Ax kernel and only the element-boundary "
               "rows gathered (bit-identical to the two-pass Ax + gather; slower on the 48^3 N=7 box: its "
               "scattered 8-byte stores cost more than the E-vector round trip they save)",
               {"TRUE", "FALSE"});
  s.newSetting("OPERATOR FACE LAYOUT", "SLOT",
               "Split operator: element-boundary results stored in the gather's row order (SLOT: the Ax kernel "
               "scatters, the gather reads contiguous copies) or in element order (ELEMENT: the gather reads "
               "through column ids)",
               {"SLOT", "ELEMENT"});
  s.newSetting("MULTIGRID COARSE SINGLE REDUCTION", "TRUE",
               "Fused coarse Jacobi-PCG in its single-reduction (Chronopoulos-Gear) form: two launches per "
               "iteration instead of three",
               {"TRUE", "FALSE"});
  s.newSetting("MULTIGRID COARSE ASSEMBLED", "TRUE",
               "Single-reduction coarse Jacobi-PCG on the assembled N=1 operator (CSR): one launch per iteration",
               {"TRUE", "FALSE"});
  s.newSetting("MULTIGRID COARSE SOLVER", "JACOBI-PCG",
               "Base level of the p-MG: FP64 Jacobi-PCG to MULTIGRID COARSE TOLERANCE, or one V-cycle of "
               "smoothed-aggregation AMG on the assembled operator (the reference's parAlmond)",
               {"JACOBI-PCG", "AMG"});
  s.newSetting("MULTIGRID COARSE CHECK EVERY", "8",
               "Coarse Jacobi-PCG: iterations queued between host convergence reads (same iterates as 1)");
  s.newSetting("MULTIGRID LEARNED WEIGHTS", "", "Weight file of the learned fine smoother");
  s.newSetting("MULTIGRID LEARNED MODE", "REPLACE",
               "REPLACE: network pre- and post-smoothing; PRE: network pre-smoothing, Chebyshev post-smoothing; "
               "AUGMENT: Chebyshev smoothing plus one network step after the pre-smoothing",
               {"REPLACE", "PRE", "AUGMENT"});
  s.newSetting("MULTIGRID LEARNED ARCH", "MLP",
               "MLP: 896-256-256-512 ReLU network with the energy-optimal step; LINEAR: one bias-free linear layer "
               "(a learned element-overlapping local solve) with the fixed step, a linear smoother",
               {"MLP", "LINEAR"});
  s.newSetting("MULTIGRID LEARNED SWEEPS", "1", "Applications of the linear learned smoother per smoothing");
  s.newSetting("LINEAR SOLVER TOLERANCE", "1e-8");
  s.newSetting("MAXIMUM ITERATIONS", "5000");
  s.newSetting("DATA SEED", "1", "Seed of the Fourier forcing data set");
  return s;
}

platform_t::platform_t(const settings_t& s) {
  int dev = 0;
  s.getSetting("DEVICE NUMBER", dev);
  int ndev = 0;
  LIBP_HIP(hipGetDeviceCount(&ndev));
  LIBP_ABORT("No HIP device available", ndev <= 0);
  LIBP_ABORT("DEVICE NUMBER " << dev << " out of range", dev < 0 || dev >= ndev);
  device = dev;
  LIBP_HIP(hipSetDevice(device));
  hipStream_t st;
  LIBP_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  double* pinned = nullptr;
  LIBP_HIP(hipHostMalloc((void**)&pinned, 16 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
  hipEvent_t ev = nullptr;
  LIBP_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  struct State {
    hipStream_t st = nullptr;
    double* pinned = nullptr;
    hipEvent_t ev = nullptr;
    ~State() {
      if (ev) (void)hipEventDestroy(ev);
      if (st) (void)hipStreamDestroy(st);
      if (pinned) (void)hipHostFree(pinned);
    }
  };
  auto stt = std::make_shared<State>();
  stt->st = st;
  stt->pinned = pinned;
  stt->ev = ev;
  state = stt;
  stream = st;
  h_scalars = pinned;
  ev_read = ev;
  LIBP_HIP(hipHostGetDevicePointer((void**)&d_hScalars, pinned, 0));
  o_reduceWS = deviceMemory<char>::alloc((size_t)libp_reduce_ws_bytes());
  LIBP_HIP(hipMemsetAsync(o_reduceWS.ptr(), 0, o_reduceWS.length(), stream));
  o_scalars = deviceMemory<double>::alloc(64);
  LIBP_HIP(hipMemsetAsync(o_scalars.ptr(), 0, 64 * sizeof(double), stream));
  finish();
}

const double* platform_t::readScalars(const double* dptr, int count) const {
  LIBP_ABORT("readScalars: at most 16 values", count > 16);
  LIBP_CALL(libp_scalars_to_host_f64(count, dptr, d_hScalars, stream));
  LIBP_HIP(hipEventRecord(ev_read, stream));
  LIBP_HIP(hipEventSynchronize(ev_read));
  return h_scalars;
}

}  // namespace libp
