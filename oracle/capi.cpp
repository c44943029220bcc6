// CPU oracle (test infrastructure only): C entry points for the Python tests
// and for bench.py's CPU baseline. Nothing in the product links this library.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include <omp.h>

#include "oracle.hpp"

namespace oracle {
std::vector<std::complex<double>> hessenbergEigs(int n, const std::vector<double>& Hcm);
}

using namespace oracle;

namespace {
struct Handle {
  Settings s;
  GlibcRand rng;
  Drand48 drng;
  Mesh mesh;
  Elliptic el;
  std::unique_ptr<MultiGrid> mg;
  std::unique_ptr<AmgO> amg;  // parAlmond AMG on the handle's own assembled operator
  std::vector<double> invDiag;
  std::vector<double> scratch;
  std::string err;
};
thread_local std::string g_err;

template <typename T>
long copy_out(const std::vector<T>& v, void* out) {
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(T));
  return (long)v.size();
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_create(const char* rc_text) {
  try {
    Handle* h = new Handle();
    h->s = Settings::defaults();
    Settings user;
    user.parseText(rc_text);
    for (auto& k : user.order) h->s.set(k, user.get(k));
    h->mesh.Setup(h->s);
    h->el.Setup(h->mesh, h->s, h->s.getd("LAMBDA"), h->rng);
    return h;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_free(void* p) { delete (Handle*)p; }

long orc_geti(void* p, const char* name) {
  Handle* h = (Handle*)p;
  std::string n(name);
  if (n == "Nelements") return h->mesh.Nelements;
  if (n == "N") return h->mesh.N;
  if (n == "Np") return h->mesh.Np;
  if (n == "Nq") return h->mesh.Nq;
  if (n == "Ngather") return h->el.Ndofs;
  if (n == "Nmasked") return h->el.Nmasked;
  if (n == "NrowBlocksT") return h->el.ogsMasked.gatherLocal.NrowBlocksT;
  if (n == "NrowBlocksN") return h->el.ogsMasked.gatherLocal.NrowBlocksN;
  if (n == "allNeumann") return h->el.allNeumann;
  return -1;
}

double orc_getd(void* p, const char* name) {
  Handle* h = (Handle*)p;
  std::string n(name);
  if (n == "lambda") return h->el.lambda;
  if (n == "allNeumannScale") return h->el.allNeumannScale;
  return NAN;
}

// Copies the named array into out (if non-null) and returns its length.
// dtype: 0 float64, 1 int32, 2 int64.
long orc_array(void* p, const char* name, void* out, int* dtype) {
  Handle* h = (Handle*)p;
  std::string n(name);
  const Mesh& m = h->mesh;
  const Elliptic& e = h->el;
  const OgsOperator& G = e.ogsMasked.gatherLocal;
  *dtype = 0;
  if (n == "x") return copy_out(m.x, out);
  if (n == "y") return copy_out(m.y, out);
  if (n == "z") return copy_out(m.z, out);
  if (n == "ggeo") return copy_out(m.ggeo, out);
  if (n == "wJ") return copy_out(m.wJ, out);
  if (n == "D") return copy_out(m.D, out);
  if (n == "gllz") return copy_out(m.gllz, out);
  if (n == "gllw") return copy_out(m.gllw, out);
  if (n == "weightG") return copy_out(e.weightG, out);
  *dtype = 2;
  if (n == "globalIds") return copy_out(m.globalIds, out);
  if (n == "maskedGlobalIds") return copy_out(e.maskedGlobalIds, out);
  if (n == "EToV") return copy_out(m.EToV, out);
  *dtype = 1;
  if (n == "mapB") return copy_out(e.mapB, out);
  if (n == "meshMapB") return copy_out(m.mapB, out);
  if (n == "EToE") return copy_out(m.EToE, out);
  if (n == "EToB") return copy_out(m.EToB, out);
  if (n == "vmapM") return copy_out(m.vmapM, out);
  if (n == "vmapP") return copy_out(m.vmapP, out);
  if (n == "GlobalToLocal") return copy_out(e.GlobalToLocal, out);
  if (n == "rowStartsT") return copy_out(G.rowStartsT, out);
  if (n == "colIdsT") return copy_out(G.colIdsT, out);
  if (n == "rowStartsN") return copy_out(G.rowStartsN, out);
  if (n == "colIdsN") return copy_out(G.colIdsN, out);
  if (n == "blockRowStartsT") return copy_out(G.blockRowStartsT, out);
  if (n == "blockRowStartsN") return copy_out(G.blockRowStartsN, out);
  if (n == "faceNodes") return copy_out(m.faceNodes, out);
  *dtype = -1;
  return -1;
}

void orc_partial_ax(void* p, const double* q, double* AqL) { ((Handle*)p)->el.PartialAx(q, AqL); }

void orc_partial_ax_omp(void* p, const double* q, double* AqL) { ((Handle*)p)->el.PartialAx(q, AqL, true); }

int orc_omp_threads() {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}

void orc_operator(void* p, const double* q, double* Aq) {
  Handle* h = (Handle*)p;
  h->el.Operator(q, Aq, h->scratch);
}

void orc_gather(void* p, const double* v, double* gv, int trans) {
  ((Handle*)p)->el.ogsMasked.gatherLocal.GatherAdd(gv, v, trans != 0);
}

void orc_scatter(void* p, const double* gv, double* v) {
  ((Handle*)p)->el.ogsMasked.gatherLocal.Scatter(v, gv, false);
}

void orc_diagonal(void* p, double* diag) {
  std::vector<double> d;
  ((Handle*)p)->el.BuildOperatorDiagonal(d);
  std::memcpy(diag, d.data(), d.size() * sizeof(double));
}

void orc_rhs_local(void* p, double* rL) {
  std::vector<double> r;
  ((Handle*)p)->el.Rhs(r);
  std::memcpy(rL, r.data(), r.size() * sizeof(double));
}

double orc_inner(long N, const double* x, const double* y) { return linalg_innerProd((dlong)N, x, y); }

// Jacobi-PCG (or PCG without preconditioner when jacobi==0) on the handle's
// operator. Returns the iteration count; x and r are updated in place.
int orc_pcg(void* p, double* x, double* r, double tol, int maxit, int jacobi, int pureRelative, double* rdotr_final) {
  Handle* h = (Handle*)p;
  struct A_ : OpFn {
    Handle* h;
    void apply(const double* a, double* b) override { h->el.Operator(a, b, h->scratch); }
  } A;
  A.h = h;
  struct M_ : OpFn {
    std::vector<double>* inv;
    dlong N;
    void apply(const double* a, double* b) override {
      if (inv)
        for (dlong n = 0; n < N; ++n) b[n] = 1.0 * (*inv)[n] * a[n];
      else
        for (dlong n = 0; n < N; ++n) b[n] = a[n];
    }
  } M;
  M.N = h->el.Ndofs;
  M.inv = nullptr;
  if (jacobi) {
    std::vector<double> d;
    h->el.BuildOperatorDiagonal(d);
    h->invDiag.resize(d.size());
    for (size_t n = 0; n < d.size(); ++n) h->invDiag[n] = 1.0 / d[n];
    M.inv = &h->invDiag;
  }
  PCGResult res = pcg_solve(h->el.Ndofs, A, M, x, r, tol, maxit, false, false, pureRelative != 0);
  if (rdotr_final) *rdotr_final = res.rdotrFinal;
  return res.iters;
}

// Multigrid preconditioner on the handle (degrees from the settings).
int orc_mg_setup(void* p) {
  Handle* h = (Handle*)p;
  try {
    h->mg.reset(new MultiGrid());
    if (h->s.has("MULTIGRID COARSE TOLERANCE")) h->mg->coarseTol = h->s.getd("MULTIGRID COARSE TOLERANCE");
    if (h->s.has("MULTIGRID COARSE MAXIT")) h->mg->coarseMaxIt = h->s.geti("MULTIGRID COARSE MAXIT");
    // same schedule logic as elliptic_run
    std::vector<int> deg;
    if (h->s.compare("MULTIGRID COARSENING", "CUSTOM")) {
      std::string v = h->s.get("MULTIGRID DEGREES");
      size_t a = 0;
      while (a < v.size()) {
        size_t b = v.find(',', a);
        if (b == std::string::npos) b = v.size();
        deg.push_back(std::stoi(v.substr(a, b - a)));
        a = b + 1;
      }
    } else {
      throw std::runtime_error("orc_mg_setup expects MULTIGRID COARSENING CUSTOM");
    }
    h->mg->Setup(h->el, h->rng, h->drng, deg);
    return (int)h->mg->levels.size();
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void orc_mg_apply(void* p, const double* r, double* z) { ((Handle*)p)->mg->apply(r, z); }

// Gathered length of level l of the handle's multigrid (-1 if there is no such level).
long orc_mg_level_ndofs(void* p, int l) {
  Handle* h = (Handle*)p;
  if (!h->mg || l < 0 || l >= (int)h->mg->levels.size()) return -1;
  return h->mg->levels[l]->Ndofs;
}

// One operation of p-multigrid level l (0 <= l < number of levels - 1):
//   op 0: out = A_l in;  op 1: out = R in (coarsen l -> l+1);
//   op 2: out += P in (prolongate l+1 -> l);  op 3: smoother, in = rhs, out = x
//   (updated in place, flag = x is zero);  op 4: out = invDiag of level l.
int orc_mg_level_op(void* p, int l, int op, const double* in, double* out, int flag) {
  Handle* h = (Handle*)p;
  if (!h->mg || l < 0 || l + 1 >= (int)h->mg->levels.size()) {
    g_err = "no such multigrid level";
    return -1;
  }
  const MGLevelO& L = h->mg->mg[l];
  switch (op) {
    case 0: {
      std::vector<double> scratch;
      L.el->Operator(in, out, scratch);
      return 0;
    }
    case 1: coarsen(L, in, out); return 0;
    case 2: prolongate(L, in, out); return 0;
    case 3:
      if (L.stype == 1) smoothChebyshev(L, in, out, flag != 0);
      else smoothJacobi(L, in, out, flag != 0);
      return 0;
    case 4: std::copy(L.invDiagA.begin(), L.invDiagA.end(), out); return 0;
    default: g_err = "unknown multigrid level operation"; return -1;
  }
}

// rho(D^-1 A) of the handle's own operator by the smoother setup's 10-step
// Arnoldi, started from the drand48 stream seeded with `seed`.
double orc_max_eig(void* p, long seed) {
  Handle* h = (Handle*)p;
  std::vector<double> d;
  h->el.BuildOperatorDiagonal(d);
  for (auto& v : d) v = 1.0 / v;
  Drand48 r;
  r.seed(seed);
  return maxEigSmoothAx(h->el, d, r);
}

/* ---- parAlmond AMG ---- */
namespace {
// which = 0: the AMG on the handle's own operator (orc_amg_setup); 1: the p-MG's base level
const AmgO* amgOf(Handle* h, int which) {
  if (which == 0) return h->amg.get();
  return h->mg ? h->mg->amg.get() : nullptr;
}
}  // namespace

// Assembled operator (BuildOperatorMatrixContinuousHex3D) of the handle: returns nnz;
// fills rowStarts [Ngather+1], cols, vals when non-null.
long orc_assembled_csr(void* p, int* rs, int* ci, double* v) {
  Handle* h = (Handle*)p;
  try {
    CsrO A = BuildOperatorMatrixContinuousHex3D(h->el);
    if (rs) std::copy(A.rs.begin(), A.rs.end(), rs);
    if (ci) std::copy(A.ci.begin(), A.ci.end(), ci);
    if (v) std::copy(A.v.begin(), A.v.end(), v);
    return A.nnz();
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// AMG setup on the handle's own assembled operator, drand48 seeded with `seed`.
// Returns the number of levels (-1 on error).
int orc_amg_setup(void* p, long seed) {
  Handle* h = (Handle*)p;
  try {
    Drand48 d;
    d.seed(seed);
    h->amg.reset(new AmgO());
    h->amg->Setup(BuildOperatorMatrixContinuousHex3D(h->el), h->s, d, h->el.allNeumann != 0,
                  h->el.allNeumannPenalty);
    return (int)h->amg->levels.size();
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int orc_amg_levels(void* p, int which) {
  const AmgO* a = amgOf((Handle*)p, which);
  return a ? (int)a->levels.size() : -1;
}

double orc_amg_level_d(void* p, int which, int l, const char* name) {
  const AmgO* a = amgOf((Handle*)p, which);
  if (!a || l < 0 || l >= (int)a->levels.size()) return NAN;
  const AmgLevelO& L = a->levels[l];
  std::string n(name);
  if (n == "rho") return L.rho;
  if (n == "lambda") return L.lambda;
  if (n == "lambda0") return L.lambda0;
  if (n == "lambda1") return L.lambda1;
  return NAN;
}

// Named array of AMG level l; dtype 0 float64, 1 int32, 2 int64. Returns the length.
long orc_amg_level_array(void* p, int which, int l, const char* name, void* out, int* dtype) {
  const AmgO* a = amgOf((Handle*)p, which);
  if (!a || l < 0 || l >= (int)a->levels.size()) return -1;
  const AmgLevelO& L = a->levels[l];
  std::string n(name);
  *dtype = 0;
  if (n == "A_v") return copy_out(L.A.v, out);
  if (n == "P_v") return copy_out(L.P.v, out);
  if (n == "R_v") return copy_out(L.R.v, out);
  if (n == "diagInv") return copy_out(L.diagInv, out);
  if (n == "coarseInv") return copy_out(a->coarseInv, out);
  if (n == "weights") return copy_out(L.weights, out);
  if (n == "null") return copy_out(a->null, out);
  *dtype = 1;
  if (n == "states") return copy_out(L.states, out);
  if (n == "A_rs") return copy_out(L.A.rs, out);
  if (n == "A_ci") return copy_out(L.A.ci, out);
  if (n == "P_rs") return copy_out(L.P.rs, out);
  if (n == "P_ci") return copy_out(L.P.ci, out);
  if (n == "R_rs") return copy_out(L.R.rs, out);
  if (n == "R_ci") return copy_out(L.R.ci, out);
  if (n == "C_rs") return copy_out(L.Crs, out);
  if (n == "C_ci") return copy_out(L.Cci, out);
  *dtype = 2;
  if (n == "FineToCoarse") return copy_out(L.FineToCoarse, out);
  *dtype = -1;
  return -1;
}

// x = V(b), one AMG V-cycle from level 0 (x need not be initialised)
int orc_amg_vcycle(void* p, int which, const double* b, double* x) {
  const AmgO* a = amgOf((Handle*)p, which);
  if (!a) {
    g_err = "no AMG";
    return -1;
  }
  a->vcycle(0, b, x);
  return 0;
}

// one smoothing of AMG level l (x updated in place; flag = x is zero)
int orc_amg_smooth(void* p, int which, int l, const double* b, double* x, int flag) {
  const AmgO* a = amgOf((Handle*)p, which);
  if (!a || l < 0 || l + 1 >= (int)a->levels.size()) {
    g_err = "no such AMG level";
    return -1;
  }
  a->smooth(l, b, x, flag != 0);
  return 0;
}

double orc_mg_level_d(void* p, int l, const char* name) {
  Handle* h = (Handle*)p;
  std::string n(name);
  if (n == "lambda1") return h->mg->mg[l].lambda1;
  if (n == "lambda0") return h->mg->mg[l].lambda0;
  return NAN;
}

// End-to-end run of a settings text. Returns the solution norm (NaN on error).
double orc_run(const char* rc_text, int* iters, long* ngather) {
  try {
    Settings s = Settings::defaults();
    Settings user;
    user.parseText(rc_text);
    for (auto& k : user.order) s.set(k, user.get(k));
    RunResult r = elliptic_run(s, 0);
    if (iters) *iters = r.iters;
    if (ngather) *ngather = r.Ngather;
    return r.norm;
  } catch (std::exception& e) {
    g_err = e.what();
    return NAN;
  }
}

void orc_gll(int N, double* x, double* w) {
  std::vector<double> xx, ww;
  JacobiGLL(N, xx, ww);
  std::memcpy(x, xx.data(), sizeof(double) * (N + 1));
  std::memcpy(w, ww.data(), sizeof(double) * (N + 1));
}

void orc_dmatrix(int N, double* D) {
  std::vector<double> x, w, d;
  JacobiGLL(N, x, w);
  Dmatrix1D(N, x, d);
  std::memcpy(D, d.data(), sizeof(double) * d.size());
}

void orc_interp(int Nc, int Nf, double* P) {
  std::vector<double> p;
  DegreeRaiseMatrix1D(Nc, Nf, p);
  std::memcpy(P, p.data(), sizeof(double) * p.size());
}

void orc_rand_seq(int n, int* out) {
  GlibcRand r;
  for (int i = 0; i < n; ++i) out[i] = r.next();
}

void orc_drand48_seq(long seed, int n, double* out) {
  Drand48 d;
  d.seed(seed);
  for (int i = 0; i < n; ++i) out[i] = d.next();
}

int orc_hess_eigs(int n, const double* Hcm, double* re, double* im) {
  try {
    std::vector<double> H(Hcm, Hcm + n * n);
    auto ev = hessenbergEigs(n, H);
    for (int i = 0; i < (int)ev.size(); ++i) {
      re[i] = ev[i].real();
      im[i] = ev[i].imag();
    }
    return (int)ev.size();
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// CPU baseline: time `reps` applications of the partial Ax on the handle's
// mesh. Returns the median seconds per application.
double orc_time_partial_ax(void* p, int reps, int parallel) {
  Handle* h = (Handle*)p;
  const size_t NL = (size_t)h->mesh.Nelements * h->mesh.Np;
  std::vector<double> q(h->el.Ndofs), AqL(NL);
  for (size_t i = 0; i < q.size(); ++i) q[i] = std::sin(0.37 * (double)i);
  std::vector<double> t;
  for (int r = 0; r < reps; ++r) {
    auto a = std::chrono::steady_clock::now();
    h->el.PartialAx(q.data(), AqL.data(), parallel != 0);
    auto b = std::chrono::steady_clock::now();
    t.push_back(std::chrono::duration<double>(b - a).count());
  }
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

// CPU baseline legs on the handle's mesh: leg 0 the gather (Trans) of an
// E-vector, 1 the whole operator (partial Ax + gather), 2 a Jacobi-PCG solve of
// the handle's right-hand side with the reference's stopping rule. Returns the
// median seconds of `reps` runs; *iters receives the solve's iteration count.
double orc_time_leg(void* p, int leg, int reps, int parallel, int* iters) {
  Handle* h = (Handle*)p;
  Elliptic& el = h->el;
  const size_t NL = (size_t)h->mesh.Nelements * h->mesh.Np;
  el.parallel = parallel != 0;
  std::vector<double> q(el.Ndofs), Aq(el.Ndofs), vL(NL), t;
  for (size_t i = 0; i < q.size(); ++i) q[i] = std::sin(0.37 * (double)i);
  for (size_t i = 0; i < NL; ++i) vL[i] = std::cos(0.11 * (double)i);
  std::vector<double> rL, r0(el.Ndofs), diagA, invD;
  if (leg == 2) {
    el.Rhs(rL);
    el.ogsMasked.gatherLocal.GatherAdd(r0.data(), rL.data(), true);
    el.BuildOperatorDiagonal(diagA);
    invD.resize(diagA.size());
    for (size_t n = 0; n < diagA.size(); ++n) invD[n] = 1.0 / diagA[n];
  }
  struct A_ : OpFn {
    Elliptic* el;
    std::vector<double> s;
    void apply(const double* a, double* b) override { el->Operator(a, b, s); }
  } A;
  A.el = &el;
  struct M_ : OpFn {
    const std::vector<double>* inv;
    void apply(const double* a, double* b) override {
      for (size_t n = 0; n < inv->size(); ++n) b[n] = 1.0 * (*inv)[n] * a[n];
    }
  } M;
  M.inv = &invD;
  for (int r = 0; r < reps; ++r) {
    auto a = std::chrono::steady_clock::now();
    if (leg == 0) {
      el.ogsMasked.gatherLocal.GatherAdd(Aq.data(), vL.data(), true, parallel != 0);
    } else if (leg == 1) {
      el.Operator(q.data(), Aq.data(), h->scratch);
    } else {
      std::vector<double> x(el.Ndofs, 0.0), rr = r0;
      PCGResult res = pcg_solve(el.Ndofs, A, M, x.data(), rr.data(), 1e-8, 5000, false, false, false);
      if (iters) *iters = res.iters;
    }
    auto b = std::chrono::steady_clock::now();
    t.push_back(std::chrono::duration<double>(b - a).count());
  }
  el.parallel = false;
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

/* ---- learned smoother oracle ---- */
long orc_learned_trace(void* p, int* out) {
  Handle* h = (Handle*)p;
  std::vector<int> t;
  LearnedTraceIds(h->el, t);
  if (out) std::memcpy(out, t.data(), sizeof(int) * t.size());
  return (long)t.size();
}

double orc_learned_loss_grad(void* p, int H, const double* theta, const double* r, const double* eh, double nrm,
                             const double* invD, int bf, double* grad, double* c) {
  Handle* h = (Handle*)p;
  std::vector<int> t;
  LearnedTraceIds(h->el, t);
  return LearnedLossGrad(h->el, t, H, theta, r, eh, nrm, invD, bf != 0, grad, c);
}

void orc_learned_forward(void* p, int H, const double* theta, const double* r, const double* invD, int bf, double* X,
                         double* scale, double* Y, double* c) {
  Handle* h = (Handle*)p;
  std::vector<int> t;
  LearnedTraceIds(h->el, t);
  LearnedForward(h->el, t, H, theta, r, invD, bf != 0, X, scale, Y, c);
}

void orc_learned_linear_gram(void* p, const double* r, const double* eh, const double* invD, int bf, double* G,
                             double* B) {
  Handle* h = (Handle*)p;
  std::vector<int> t;
  LearnedTraceIds(h->el, t);
  LearnedLinearGram(h->el, t, r, eh, invD, bf != 0, G, B);
}

double orc_bf16r(double x) { return bf16r(x); }

}  // extern "C"
