// C ABI over elliptic_t: lets Python (ctypes), bench drivers and foreign
// callers build and drive the C++ solver without C++ types at the boundary.
#include <cmath>
#include <cstring>
#include <string>

#include <sstream>

#include "amg.hpp"
#include "elliptic.hpp"
#include "learned.hpp"

namespace libp_amd {
void set_last_error(const char* where, const char* what);
}

using namespace libp;

namespace {
struct Handle {
  settings_t settings;
  platform_t platform;
  mesh_t mesh;
  std::shared_ptr<elliptic_t> elliptic;
  ellipticRunResult_t last;
  std::shared_ptr<csrHost_t> csr;    // assembled fine operator (test hook)
  std::shared_ptr<amgSolver_t> amg;  // AMG on the fine operator (test hook)
};

template <typename F>
int guarded(const char* where, F f) {
  try {
    f();
    return LIBP_OK;
  } catch (std::exception& e) {
    libp_amd::set_last_error(where, e.what());
    return LIBP_ERR_SETUP;
  }
}
}  // namespace

extern "C" void* libp_elliptic_create(const char* rc_text) {
  Handle* h = nullptr;
  int st = guarded("libp_elliptic_create", [&]() {
    LIBP_ABORT("null settings text", rc_text == nullptr);
    std::unique_ptr<Handle> hh(new Handle());
    hh->settings = ellipticDefaultSettings();
    hh->settings.readSettingsFromText(rc_text);
    hh->platform = platform_t(hh->settings);
    hh->mesh.Setup(hh->platform, hh->settings, comm_t());
    double lambda = 0.0;
    hh->settings.getSetting("LAMBDA", lambda);
    memory<int> BCType(3);
    BCType[0] = 0;
    BCType[1] = 1;
    BCType[2] = 2;
    hh->elliptic = std::make_shared<elliptic_t>(hh->platform, hh->mesh, hh->settings, lambda, 3, BCType);
    hh->platform.finish();
    h = hh.release();
  });
  return st == LIBP_OK ? h : nullptr;
}

extern "C" void libp_elliptic_free(void* p) { delete (Handle*)p; }

extern "C" long libp_elliptic_geti(void* p, const char* name) {
  Handle* h = (Handle*)p;
  if (!h || !name) return -1;
  const std::string n(name);
  elliptic_t& e = *h->elliptic;
  if (n == "Nelements") return e.mesh.Nelements;
  if (n == "N") return e.mesh.N;
  if (n == "Np") return e.mesh.Np;
  if (n == "Nq") return e.mesh.Nq;
  if (n == "Ngather") return e.Ndofs;
  if (n == "Nmasked") return e.Nmasked;
  if (n == "allNeumann") return e.allNeumann;
  if (n == "NrowBlocksT") return e.ogsMasked.NrowBlocksT;
  if (n == "NrowBlocksN") return e.ogsMasked.NrowBlocksN;
  if (n == "lastIterations") return h->last.iterations;
  if (n == "splitOperator") return e.SplitGather() ? 1 : 0;
  if (n == "faceRows") return e.SplitGather() ? e.SplitGather()->NfaceRows : -1;
  if (n == "coarseIterations") {
    if (auto* mg = dynamic_cast<MultiGridPrecon*>(e.precon.get())) return multigridCoarseIterations(*mg);
    return -1;
  }
  return -1;
}

extern "C" double libp_elliptic_getd(void* p, const char* name) {
  Handle* h = (Handle*)p;
  if (!h || !name) return NAN;
  const std::string n(name);
  elliptic_t& e = *h->elliptic;
  if (n == "lambda") return e.lambda;
  if (n == "allNeumannScale") return e.allNeumannScale;
  if (n == "curveAmp") return e.mesh.curveAmp;
  return NAN;
}

extern "C" void* libp_elliptic_array(void* p, const char* name, long* length, int* dtype, int* on_host) {
  Handle* h = (Handle*)p;
  if (!h || !name) return nullptr;
  const std::string n(name);
  elliptic_t& e = *h->elliptic;
  mesh_t& m = e.mesh;
  *on_host = 0;
  auto dev = [&](auto& mem, int dt) -> void* {
    *length = (long)mem.length();
    *dtype = dt;
    return (void*)mem.ptr();
  };
  auto host = [&](auto& mem, int dt) -> void* {
    *length = (long)mem.length();
    *dtype = dt;
    *on_host = 1;
    return (void*)mem.ptr();
  };
  if (n == "ggeo") return dev(m.o_ggeo, 0);
  if (n == "wJ") return dev(m.o_wJ, 0);
  if (n == "x") return dev(m.o_x, 0);
  if (n == "y") return dev(m.o_y, 0);
  if (n == "z") return dev(m.o_z, 0);
  if (n == "D") return dev(m.o_D, 0);
  if (n == "gllz") return dev(m.o_gllz, 0);
  if (n == "gllw") return dev(m.o_gllw, 0);
  if (n == "weightG") return dev(e.o_weightG, 0);
  if (n == "GlobalToLocal") return dev(e.o_GlobalToLocal, 1);
  if (n == "rowStartsT") return dev(e.ogsMasked.o_rowStartsT, 1);
  if (n == "colIdsT") return dev(e.ogsMasked.o_colIdsT, 1);
  if (n == "rowStartsN") return dev(e.ogsMasked.o_rowStartsN, 1);
  if (n == "colIdsN") return dev(e.ogsMasked.o_colIdsN, 1);
  if (n == "mapBDevice") return dev(e.o_mapB, 1);
  if (n == "invDiag") {
    if (auto* j = dynamic_cast<JacobiPrecon*>(e.precon.get())) return dev(j->o_invDiagA, 0);
    return nullptr;
  }
  if (auto* mg = dynamic_cast<MultiGridPrecon*>(e.precon.get())) {
    static thread_local deviceMemory<float> keep;
    if (n == "learnedInvDiag") {
      keep = multigridFineInvDiag32(*mg);
      return keep.ptr() ? dev(keep, 3) : nullptr;
    }
    if (learnedSmoother_t* L = multigridLearned(*mg)) {
      if (n == "learnedTheta") return dev(L->o_theta, 3);
      if (n == "learnedTrace") return dev(L->o_trace, 1);
      if (n == "learnedGrad") return dev(L->o_grad, 3);
      if (n == "learnedScale") return dev(L->o_scale, 3);
      if (n == "learnedY") return dev(L->o_Y, 3);
    }
  }
  if (n == "globalIds") return host(m.globalIds, 2);
  if (n == "maskedGlobalIds") return host(e.maskedGlobalIds, 2);
  if (n == "meshMapB") return host(m.mapB, 1);
  if (n == "mapB") return host(e.mapB, 1);
  if (n == "EToV") return host(m.EToV, 2);
  if (n == "blockRowStartsT") return host(e.ogsMasked.blockRowStartsT, 1);
  if (n == "blockRowStartsN") return host(e.ogsMasked.blockRowStartsN, 1);
  return nullptr;
}

extern "C" void* libp_elliptic_stream(void* p) {
  Handle* h = (Handle*)p;
  return h ? (void*)h->platform.stream : nullptr;
}

extern "C" int libp_elliptic_operator(void* p, const double* q, double* Aq) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_operator", [&]() {
    LIBP_ABORT("null handle/pointer", !h || !q || !Aq);
    const dlong N = h->elliptic->Ndofs;
    deviceMemory<dfloat> o_q = deviceMemory<dfloat>::wrap(const_cast<double*>(q), N);
    deviceMemory<dfloat> o_Aq = deviceMemory<dfloat>::wrap(Aq, N);
    h->elliptic->Operator(o_q, o_Aq);
  });
}

extern "C" int libp_elliptic_partial_ax(void* p, const double* q, double* AqL) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_partial_ax", [&]() {
    LIBP_ABORT("null handle/pointer", !h || !q || !AqL);
    elliptic_t& e = *h->elliptic;
    deviceMemory<dfloat> o_q = deviceMemory<dfloat>::wrap(const_cast<double*>(q), e.Ndofs);
    deviceMemory<dfloat> o_AqL = deviceMemory<dfloat>::wrap(AqL, (size_t)e.mesh.Nelements * e.mesh.Np);
    e.PartialAx(o_q, o_AqL);
  });
}

extern "C" int libp_elliptic_precon(void* p, const double* r, double* z) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_precon", [&]() {
    LIBP_ABORT("null handle/pointer", !h || !r || !z);
    const dlong N = h->elliptic->Ndofs;
    deviceMemory<dfloat> o_r = deviceMemory<dfloat>::wrap(const_cast<double*>(r), N);
    deviceMemory<dfloat> o_z = deviceMemory<dfloat>::wrap(z, N);
    h->elliptic->precon.Operator(o_r, o_z);
  });
}

extern "C" int libp_elliptic_run(void* p, double* out, int nout) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_run", [&]() {
    LIBP_ABORT("null handle", !h);
    h->last = h->elliptic->Run();
    const double v[6] = {h->last.norm, (double)h->last.iterations, h->last.solveSeconds,
                         h->last.rFinal, h->last.r0, h->last.maxError};
    for (int i = 0; i < nout && i < 6; ++i) out[i] = v[i];
  });
}

extern "C" int libp_elliptic_solution(void* p, double* xL_host) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_solution", [&]() {
    LIBP_ABORT("no solution available", !h || !h->elliptic->o_xL_last.isInitialized());
    h->elliptic->o_xL_last.copyTo(xL_host, h->elliptic->o_xL_last.length(), 0, h->platform.stream);
    h->platform.finish();
  });
}

/* ------------------------- p-MG level access (tests) ------------------------- */
namespace {
MultiGridPrecon& multigridOf(Handle* h) {
  LIBP_ABORT("null handle", !h);
  auto* mg = dynamic_cast<MultiGridPrecon*>(h->elliptic->precon.get());
  LIBP_ABORT("the handle has no p-MG preconditioner", !mg);
  return *mg;
}
}  // namespace

extern "C" long libp_elliptic_mg_level_ndofs(void* p, int level, int* bits) {
  long n = -1;
  int st = guarded("libp_elliptic_mg_level_ndofs", [&]() {
    MultiGridPrecon& mg = multigridOf((Handle*)p);
    if (bits) *bits = multigridPrecisionBits(mg);
    n = multigridLevelNdofs(mg, level);
  });
  return st == LIBP_OK ? n : -st;
}

extern "C" int libp_elliptic_mg_level_op(void* p, int level, int op, const void* in, void* out, int flag) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_mg_level_op", [&]() {
    LIBP_ABORT("null pointer", (!in && op != 4) || !out);
    multigridLevelOp(multigridOf(h), level, op, in, out, flag);
  });
}

extern "C" int libp_elliptic_mg_level_lambda(void* p, int level, double* lambda) {
  return guarded("libp_elliptic_mg_level_lambda", [&]() {
    LIBP_ABORT("null pointer", !lambda);
    multigridLevelLambda(multigridOf((Handle*)p), level, lambda);
  });
}

/* ------------------------- learned fine smoother ------------------------- */
namespace {
learnedSmoother_t* learnedOf(Handle* h) {
  LIBP_ABORT("null handle", !h);
  auto* mg = dynamic_cast<MultiGridPrecon*>(h->elliptic->precon.get());
  LIBP_ABORT("the handle has no p-MG preconditioner", !mg);
  learnedSmoother_t* L = multigridLearned(*mg);
  LIBP_ABORT("the handle has no learned smoother (MULTIGRID SMOOTHER LEARNED, MULTIGRID PRECISION FP32)", !L);
  return L;
}
}  // namespace

extern "C" int libp_elliptic_learned_train(void* p, const char* options, double* loss, int nloss) {
  Handle* h = (Handle*)p;
  int n = 0;
  int st = guarded("libp_elliptic_learned_train", [&]() {
    learnedOf(h);
    learnedTrainOptions opt;
    int verbose = 0;
    auto list = [](const std::string& v) {
      std::vector<long> out;
      std::stringstream ss(v);
      std::string item;
      while (std::getline(ss, item, ',')) out.push_back(std::stol(item));
      return out;
    };
    std::stringstream ss(options ? options : "");
    std::string kv;
    while (ss >> kv) {
      const size_t eq = kv.find('=');
      LIBP_ABORT("training option without '=': " << kv, eq == std::string::npos);
      const std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
      if (k == "seeds") opt.seeds = list(v);
      else if (k == "daggerSeeds") opt.daggerSeeds = list(v);
      else if (k == "steps") opt.steps = std::stoi(v);
      else if (k == "dagger") opt.dagger = std::stoi(v);
      else if (k == "daggerSteps") opt.daggerSteps = std::stoi(v);
      else if (k == "lr") opt.lr = std::stof(v);
      else if (k == "targetTol") opt.targetTol = std::stod(v);
      else if (k == "initSeed") opt.initSeed = std::stoul(v);
      else if (k == "logEvery") opt.logEvery = std::stoi(v);
      else if (k == "maxRecord") opt.maxRecordPerSolve = std::stoi(v);
      else if (k == "samples") opt.linearSamples = std::stoi(v);
      else if (k == "ridge") opt.ridge = std::stod(v);
      else if (k == "coarseTol") opt.linearCoarseTol = std::stod(v);
      else if (k == "verbose") verbose = std::stoi(v);
      else LIBP_FORCE_ABORT("unknown training option " << k);
    }
    auto* mg = dynamic_cast<MultiGridPrecon*>(h->elliptic->precon.get());
    std::vector<double> hist = multigridTrainLearned(*mg, opt, verbose);
    for (size_t i = 0; i < hist.size() && (int)i < nloss; ++i) loss[i] = hist[i];
    n = (int)hist.size();
  });
  return st == LIBP_OK ? n : -st;
}

extern "C" int libp_elliptic_learned_load(void* p, const char* path) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_learned_load", [&]() {
    learnedSmoother_t* L = learnedOf(h);
    LIBP_ABORT(L->status, !L->load(path ? path : ""));
  });
}

extern "C" int libp_elliptic_learned_save(void* p, const char* path) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_learned_save", [&]() { learnedOf(h)->save(path ? path : ""); });
}

/* out: [0] active, [1] fallback, [2] rejected steps, [3] parameters, [4] last
 * step alpha, [5] last training loss, [6] input width, [7] hidden width,
 * [8] architecture (0 MLP, 1 linear), [9] sweeps */
extern "C" int libp_elliptic_learned_info(void* p, double* out, int nout) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_learned_info", [&]() {
    learnedSmoother_t* L = learnedOf(h);
    std::vector<double> s = L->stats();
    const double v[10] = {(double)L->active, (double)L->fallback, s[3], (double)L->nParams, s[2], s[4],
                          (double)L->nIn, (double)L->H, (double)L->arch, (double)L->sweeps};
    for (int i = 0; i < nout && i < 10; ++i) out[i] = v[i];
  });
}

extern "C" const char* libp_elliptic_learned_status(void* p) {
  static thread_local std::string msg;
  Handle* h = (Handle*)p;
  int st = guarded("libp_elliptic_learned_status", [&]() { msg = learnedOf(h)->status; });
  return st == LIBP_OK ? msg.c_str() : "no learned smoother";
}

extern "C" int libp_elliptic_learned_get_params(void* p, float* theta, long n) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_learned_get_params", [&]() {
    learnedSmoother_t* L = learnedOf(h);
    LIBP_ABORT("parameter count " << n << " != " << L->nParams, (size_t)n != L->nParams || !theta);
    std::vector<float> th = L->parameters();
    std::memcpy(theta, th.data(), n * sizeof(float));
  });
}

extern "C" int libp_elliptic_learned_set_params(void* p, const float* theta, long n, int activate) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_learned_set_params", [&]() {
    learnedSmoother_t* L = learnedOf(h);
    LIBP_ABORT("parameter count " << n << " != " << L->nParams, (size_t)n != L->nParams || !theta);
    L->setParameters(std::vector<float>(theta, theta + n));
    L->active = activate != 0;
    L->fallback = false;
    L->status = activate ? "parameters set" : "parameters set (inactive)";
  });
}

/* c = c(r) for a device FP32 gathered residual r (the network correction,
 * before the energy-optimal step). */
extern "C" int libp_elliptic_learned_correction(void* p, const float* r, float* c) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_learned_correction", [&]() {
    learnedSmoother_t* L = learnedOf(h);
    LIBP_ABORT("null pointer", !r || !c);
    auto* mg = dynamic_cast<MultiGridPrecon*>(h->elliptic->precon.get());
    deviceMemory<float> iD = multigridFineInvDiag32(*mg);
    L->correction(r, iD.ptr());
    LIBP_HIP(hipMemcpyAsync(c, L->o_c.ptr(), L->Ndofs * sizeof(float), hipMemcpyDeviceToDevice, h->platform.stream));
    h->platform.finish();
  });
}

/* ------------------------------ assembled operator / AMG ------------------------------ */
extern "C" long libp_elliptic_csr_nnz(void* p) {
  Handle* h = (Handle*)p;
  long n = 0;
  int st = guarded("libp_elliptic_csr_nnz", [&]() {
    if (!h->csr) h->csr = std::make_shared<csrHost_t>(ellipticAssembleCSR(*h->elliptic));
    n = h->csr->nnz();
  });
  return st == LIBP_OK ? n : -st;
}

extern "C" int libp_elliptic_csr(void* p, int* rowStarts, int* cols, double* vals) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_csr", [&]() {
    LIBP_ABORT("null pointer", !rowStarts || !cols || !vals);
    if (!h->csr) h->csr = std::make_shared<csrHost_t>(ellipticAssembleCSR(*h->elliptic));
    std::memcpy(rowStarts, h->csr->rs.data(), h->csr->rs.size() * sizeof(int));
    std::memcpy(cols, h->csr->ci.data(), h->csr->ci.size() * sizeof(int));
    std::memcpy(vals, h->csr->v.data(), h->csr->v.size() * sizeof(double));
  });
}

extern "C" int libp_elliptic_amg_setup(void* p) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_amg_setup", [&]() {
    unsigned short xs[3] = {0x330E, 0, 0};
    h->amg = std::make_shared<amgSolver_t>(*h->elliptic, xs);
  });
}

extern "C" int libp_elliptic_amg_vcycle(void* p, const double* b, double* x) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_amg_vcycle", [&]() {
    LIBP_ABORT("AMG not set up", !h->amg);
    LIBP_ABORT("null pointer", !b || !x);
    h->amg->vcycle(b, x);
    h->platform.finish();
  });
}

extern "C" int libp_elliptic_amg_info(void* p, double* out, int nout) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_amg_info", [&]() {
    LIBP_ABORT("AMG not set up", !h->amg);
    std::vector<double> v{(double)h->amg->levels.size()};
    for (auto& L : h->amg->levels) {
      v.push_back((double)L.A.n);
      v.push_back((double)L.A.nnz());
    }
    for (int i = 0; i < nout && i < (int)v.size(); ++i) out[i] = v[i];
  });
}

namespace {
amgSolver_t* amgOf(Handle* h, int which) {
  LIBP_ABORT("null handle", !h);
  amgSolver_t* a = nullptr;
  if (which == 0) {
    a = h->amg.get();
  } else if (auto* mg = dynamic_cast<MultiGridPrecon*>(h->elliptic->precon.get())) {
    a = multigridBaseAmg(*mg);
  }
  LIBP_ABORT("no AMG (which = " << which << ")", !a);
  return a;
}
template <typename T>
void hostOut(const std::vector<T>& v, void* out, long* length, int dt, int* dtype) {
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(T));
  *length = (long)v.size();
  *dtype = dt;
}
}  // namespace

extern "C" int libp_elliptic_amg_level_array(void* p, int which, int level, const char* name, void* out,
                                             long* length, int* dtype) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_amg_level_array", [&]() {
    LIBP_ABORT("null pointer", !name || !length || !dtype);
    amgSolver_t* a = amgOf(h, which);
    LIBP_ABORT("no AMG level " << level, level < 0 || level >= (int)a->levels.size());
    const amgSolver_t::level_t& L = a->levels[level];
    const std::string n(name);
    if (n == "A_rs") return hostOut(L.A.rs, out, length, 1, dtype);
    if (n == "A_ci") return hostOut(L.A.ci, out, length, 1, dtype);
    if (n == "A_v") return hostOut(L.A.v, out, length, 0, dtype);
    if (n == "P_rs") return hostOut(L.P.rs, out, length, 1, dtype);
    if (n == "P_ci") return hostOut(L.P.ci, out, length, 1, dtype);
    if (n == "P_v") return hostOut(L.P.v, out, length, 0, dtype);
    if (n == "R_rs") return hostOut(L.R.rs, out, length, 1, dtype);
    if (n == "R_ci") return hostOut(L.R.ci, out, length, 1, dtype);
    if (n == "R_v") return hostOut(L.R.v, out, length, 0, dtype);
    if (n == "aggregate") return hostOut(L.aggregate, out, length, 1, dtype);
    if (n == "invDiag") return hostOut(L.invDiag, out, length, 0, dtype);
    if (n == "null") return hostOut(a->null, out, length, 0, dtype);
    if (n == "rho") return hostOut(std::vector<double>{L.rho, L.lambda0, L.lambda1, L.lambda}, out, length, 0, dtype);
    LIBP_FORCE_ABORT("no AMG array named " << n);
  });
}

extern "C" int libp_elliptic_amg_smooth(void* p, int level, const double* b, double* x, int x_is_zero) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_amg_smooth", [&]() {
    amgSolver_t* a = amgOf(h, 0);
    LIBP_ABORT("no smoother on AMG level " << level, level < 0 || level + 1 >= (int)a->levels.size());
    LIBP_ABORT("null pointer", !b || !x);
    a->smooth(level, b, x, x_is_zero != 0);
    h->platform.finish();
  });
}

/* Gram matrices of one least-squares sample of the linear smoother (device
 * FP32 r and eh; host G [nIn][nIn] = X^T X and B [Np][nIn] = T^T X). */
extern "C" int libp_elliptic_learned_linear_gram(void* p, const float* r, const float* eh, double* G, double* B) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_learned_linear_gram", [&]() {
    learnedSmoother_t* L = learnedOf(h);
    LIBP_ABORT("null pointer", !r || !eh || !G || !B);
    auto* mg = dynamic_cast<MultiGridPrecon*>(h->elliptic->precon.get());
    deviceMemory<float> iD = multigridFineInvDiag32(*mg);
    std::vector<double> g, b;
    L->linearGramAccumulate(r, eh, iD.ptr(), g, b);
    std::memcpy(G, g.data(), g.size() * sizeof(double));
    std::memcpy(B, b.data(), b.size() * sizeof(double));
  });
}

/* Loss of one sample (device FP32 r and target eh, nrm = eh.A eh) and its
 * gradient (left in the "learnedGrad" array); no parameter update. */
extern "C" int libp_elliptic_learned_loss_grad(void* p, const float* r, const float* eh, double nrm, double* loss) {
  Handle* h = (Handle*)p;
  return guarded("libp_elliptic_learned_loss_grad", [&]() {
    learnedSmoother_t* L = learnedOf(h);
    LIBP_ABORT("null pointer", !r || !eh || !loss);
    auto* mg = dynamic_cast<MultiGridPrecon*>(h->elliptic->precon.get());
    deviceMemory<float> iD = multigridFineInvDiag32(*mg);
    deviceMemory<double> o_n = h->platform.malloc<double>(1);
    o_n.copyFrom(&nrm, 1, 0, h->platform.stream);
    auto A = [mg](deviceMemory<float>& x, deviceMemory<float>& y) { multigridApplyFineOperator32(*mg, x, y); };
    L->trainStep(r, eh, o_n.ptr(), iD.ptr(), A, 0.f, false);
    *loss = L->stats()[4];
  });
}
