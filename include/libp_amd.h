/*
 * libp_amd — MI355X (gfx950) kernels for libParanumal's C0 hex elliptic path.
 *
 * Thin C ABI: plain pointers (device pointers unless stated), sizes and a HIP
 * stream passed as void* (NULL = the null stream). Every entry point returns 0
 * on success and a nonzero status otherwise (libp_status_string() names it);
 * nothing here throws. The C++ layer (elliptic_t, MGLevel, pcg, ogs_t) wraps
 * these calls and turns a nonzero status into libp::exception, which is how
 * the reference reports errors (include/utils.hpp:57-76).
 *
 * Index types follow the reference: dlong = int32 for local/gathered indices
 * (include/types.h:53-60). Element lists may be NULL, meaning elements
 * 0..Nelements-1 in order (the single-rank case, where the reference's local
 * gather list is exactly that sequence).
 */
#ifndef LIBP_AMD_H
#define LIBP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LIBP_OK = 0,
  LIBP_ERR_ARG = 1,       /* invalid size / degree / null pointer */
  LIBP_ERR_HIP = 2,       /* a HIP runtime call failed */
  LIBP_ERR_SETUP = 3,     /* setup error (settings, mesh, geometry) */
  LIBP_ERR_UNSUPPORTED = 4
};

const char* libp_status_string(int status);
/* Last error message recorded by a failing call on this thread. */
const char* libp_last_error(void);

/* ------------------------------------------------------------------ */
/* Matrix-free hex Ax                                                   */
/* ------------------------------------------------------------------ */

/* Replaces ellipticPartialAxHex3D (solvers/elliptic/okl/ellipticAxHex3D.okl:156-295),
 * launched by elliptic_t::Operator (solvers/elliptic/src/ellipticOperator.cpp:42-47,69-75,97-104).
 * For each listed element e: q_L = q[GlobalToLocal[e*Np+n]] (0 where the map is -1),
 * AqL[e*Np+n] = (sum_ab D_a^T G_ab D_b q_L + lambda*wJ*q_L)[n].
 * Layouts: ggeo [e][6][Np] (G00,G01,G02,G11,G12,G22), wJ [e][Np], D [Nq][Nq]
 * row-major with D[i*Nq+m] = l'_m(r_i). 1 <= N <= 15. */
int libp_ax_hex_partial_f64(int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                            const double* wJ, const double* ggeo, const double* D, double lambda,
                            const double* q, double* AqL, void* stream);
/* The whole C0 operator A q = Q^T A_L Q q (ellipticOperator.cpp:29-107, i.e.
 * ellipticAxHex3D.okl:156-295 followed by the ogs gather, ogsKernels.okl:86-122)
 * in a split form, N <= 14. libp_ax_hex_split_f64 stores the rows of
 * element-interior nodes (one copy each) straight into the gathered Aq and the
 * element-boundary nodes into AqFace; libp_ogs_gather_face_f64 then sums the
 * remaining rows (faceRowStarts over the NfaceRows rows faceRows, copies in
 * ascending local order). Two layouts of AqFace: with faceSlots null, boundary
 * node f of element e (f = its place among the element's Nface =
 * libp_ax_hex_face_nodes(N) boundary nodes in local order) goes to e * Nface + f
 * and faceCols lists each row's copies as such indices; with faceSlots (one int
 * per e * Nface + f) it goes to AqFace[faceSlots[...]], the copy's place in the
 * row-ordered copy list, and faceCols is null (row f's copies are AqFace[
 * faceRowStarts[f] .. faceRowStarts[f+1])). The result is bit-identical to the
 * partial Ax and the full gather. papPartials, when not null, receives libp_ax_hex_split_blocks(N,
 * Nelements) partial sums of q.Aq over the interior rows for
 * libp_pcg_gather_face_pap_f64, which gathers the boundary rows, adds their
 * p.Ap partials after those and sets scalars[PAP] and scalars[ALPHA] as
 * libp_pcg_gather_pap_f64 does; partials holds nAxPartials +
 * libp_ogs_gather_face_blocks(NfaceRows) doubles. */
/* HBM bandwidth probe (no reference counterpart; measurement only). mode 0
 * reads x[0:n] once with 8-byte non-temporal loads on 65536 blocks and writes
 * the 65536 block sums to y (n >= 65536); mode 1 copies y[0:n] = x[0:n] with
 * 16-byte non-temporal loads and stores. n even, x and y 16-byte aligned. */
int libp_hbm_probe_f64(int mode, long n, const double* x, double* y, void* stream);
int libp_ax_hex_face_nodes(int N);
int libp_ax_hex_split_blocks(int N, int Nelements);
int libp_ax_hex_split_f64(int N, int Nelements, const int* elementList, const int* GlobalToLocal, const double* wJ,
                          const double* ggeo, const double* D, double lambda, const double* q, double* Aq,
                          double* AqFace, const int* faceSlots, double* papPartials, void* stream);
int libp_ogs_gather_face_blocks(int NfaceRows);
int libp_ogs_gather_face_f64(int NfaceRows, const int* faceRows, const int* faceRowStarts, const int* faceCols,
                             const double* AqFace, double* Aq, void* stream);
int libp_pcg_gather_face_pap_f64(int NfaceRows, const int* faceRows, const int* faceRowStarts,
                                 const int* faceCols, const double* AqFace, double* Ap, const double* p,
                                 int nAxPartials, double* partials, double* scalars, void* stream);
/* FP32 variant (no reference counterpart; the reference fixes pfloat at build time,
 * include/types.h:31-45) used by the p-multigrid smoother. */
int libp_ax_hex_partial_f32(int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                            const float* wJ, const float* ggeo, const float* D, float lambda,
                            const float* q, float* AqL, void* stream);
/* E-vector in, E-vector out (ellipticAxHex3D, ellipticAxHex3D.okl:28-152). Used for
 * the Dirichlet lift of ellipticRhsBCHex3D.okl:28-289. */
int libp_ax_hex_local_f64(int N, int Nelements, const double* wJ, const double* ggeo, const double* D,
                          double lambda, const double* qL, double* AqL, void* stream);
/* N=15 variant whose 16x16 tensor contractions run on FP64 MFMA
 * (v_mfma_f64_16x16x4f64). Same arguments and result as libp_ax_hex_partial_f64. */
int libp_ax_hex_partial_mfma_n15(int Nelements, const int* elementList, const int* GlobalToLocal,
                                 const double* wJ, const double* ggeo, const double* D, double lambda,
                                 const double* q, double* AqL, void* stream);

/* ------------------------------------------------------------------ */
/* Gather-scatter (single rank, signed, unique)                         */
/* ------------------------------------------------------------------ */

/* Replaces the ogs "gather" kernel with op=Add (libs/ogs/okl/ogsKernels.okl:86-122)
 * as called by ogsOperator_t::Gather (libs/ogs/ogsOperator.cpp:176-203):
 * gv[row] = sum_{g in [rowStarts[row], rowStarts[row+1])} v[colIds[g]],
 * summed from zero in ascending g (the reference order, so results are bit-exact).
 * Pass the T (all copies) CSR for Trans and the N (owner) CSR for NoTrans. */
int libp_ogs_gather_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v, double* gv,
                        void* stream);
int libp_ogs_gather_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v, float* gv,
                        void* stream);
/* Gather fused with the step after it (no reference counterpart as one call:
 * ogsOperator.cpp:186-201 followed by linAlg.axpy, linAlg.cpp:37-224):
 * residual: gv = r - gather(v), the r - A x of an operator whose last step is
 * the gather; add: gv += gather(v). Bit-identical to the two calls. */
int libp_ogs_gather_residual_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v, const double* r,
                                 double* gv, void* stream);
int libp_ogs_gather_residual_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v, const float* r,
                                 float* gv, void* stream);
int libp_ogs_gather_add_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v, double* gv,
                            void* stream);
int libp_ogs_gather_add_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v, float* gv,
                            void* stream);
/* Gather fused with a Chebyshev step of the p-MG smoother
 * (ellipticPreconMultiGridLevel.cpp:154-206): v is the E-vector A d (start:
 * A x) before its gather. start: RES = invDiag.*(r - Av), d = invTheta*RES,
 * X += d; step: RES -= invDiag.*Av, d = c1*RES + c2*d, X += d. Bit-identical to
 * the gather followed by libp_cheb_start / libp_cheb_step. */
int libp_ogs_gather_cheb_start_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v,
                                   const double* invDiag, const double* r, double invTheta, double* RES, double* d,
                                   double* X, void* stream);
int libp_ogs_gather_cheb_start_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v,
                                   const float* invDiag, const float* r, float invTheta, float* RES, float* d, float* X,
                                   void* stream);
int libp_ogs_gather_cheb_step_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v,
                                  const double* invDiag, double c1, double c2, double* RES, double* d, double* X,
                                  void* stream);
int libp_ogs_gather_cheb_step_f32(int Nrows, const int* rowStarts, const int* colIds, const float* v,
                                  const float* invDiag, float c1, float c2, float* RES, float* d, float* X,
                                  void* stream);
/* Replaces "scatter" (ogsKernels.okl:127-164, ogsOperator.cpp:295-321):
 * v[colIds[g]] = gv[row] for every g of the row. */
int libp_ogs_scatter_f64(int Nrows, const int* rowStarts, const int* colIds, const double* gv, double* v,
                         void* stream);
int libp_ogs_scatter_i32(int Nrows, const int* rowStarts, const int* colIds, const int* gv, int* v,
                         void* stream);

/* ------------------------------------------------------------------ */
/* BLAS-1 (libs/linAlg/okl, wrapped by libs/linAlg/linAlg.cpp)           */
/* ------------------------------------------------------------------ */
/* y = alpha*x + beta*y (y not read when beta == 0) — linAlgAXPY.okl:27-39 */
int libp_axpy_f64(int N, double alpha, const double* x, double beta, double* y, void* stream);
int libp_axpy_f32(int N, float alpha, const float* x, float beta, float* y, void* stream);
/* z = alpha*x + beta*y — linAlgAXPY.okl:41-51 */
int libp_zaxpy_f64(int N, double alpha, const double* x, double beta, const double* y, double* z, void* stream);
/* x = alpha*a.*x — linAlgAMXPY.okl:27-35 */
int libp_amx_f64(int N, double alpha, const double* a, double* x, void* stream);
int libp_amx_f32(int N, float alpha, const float* a, float* x, void* stream);
/* y = alpha*a.*x + beta*y (y not read when beta == 0) — linAlgAMXPY.okl:37-50 */
int libp_amxpy_f64(int N, double alpha, const double* a, const double* x, double beta, double* y, void* stream);
int libp_amxpy_f32(int N, float alpha, const float* a, const float* x, float beta, float* y, void* stream);
/* x = alpha — linAlgSet.okl; x += alpha — linAlgAdd.okl; x *= alpha — linAlgScale.okl */
int libp_set_f64(int N, double alpha, double* x, void* stream);
/* hostDst[0, n) = src[0, n), n <= 64, written by a kernel into pinned host
 * memory that is mapped for the device (hipHostMalloc with
 * hipHostMallocMapped; hostDst is its device pointer). Replaces the host read
 * of the solver's reduction results (rdotr.copyFrom(o_rdotr) at
 * linearSolverPCG.cpp:163): the caller waits on an event recorded after it. */
int libp_scalars_to_host_f64(int n, const double* src, double* hostDst, void* stream);
int libp_add_f64(int N, double alpha, double* x, void* stream);
int libp_scale_f64(int N, double alpha, double* x, void* stream);
/* Precision conversion (no reference counterpart). */
int libp_convert_f64_to_f32(int N, const double* x, float* y, void* stream);
int libp_convert_f32_to_f64(int N, const float* x, double* y, void* stream);

/* Reductions to a device scalar. Replace innerProd1/2, norm2_1/2, sum1/2
 * (linAlgInnerProd.okl:28-100, linAlgNorm2.okl, linAlgSum.okl) plus the
 * host copy of linAlg.cpp:160-224: the result stays in device memory at
 * *result. ws is a device workspace of at least libp_reduce_ws_bytes()
 * bytes holding one partial per block; it needs no initialisation.
 * Deterministic: fixed block partition, partials summed in block order. */
long libp_reduce_ws_bytes(void);
int libp_inner_prod_f64(int N, const double* x, const double* y, double* result, void* ws, void* stream);
int libp_weighted_inner_prod_f64(int N, const double* w, const double* x, const double* y, double* result,
                                 void* ws, void* stream);
int libp_norm2sq_f64(int N, const double* x, double* result, void* ws, void* stream);
int libp_sum_f64(int N, const double* x, double* result, void* ws, void* stream);

/* ------------------------------------------------------------------ */
/* PCG / FPCG fused vector kernels (libs/linearSolver)                  */
/* ------------------------------------------------------------------ */
/* Device-resident scalar block used by the fused PCG kernels (doubles):
 *  [0] rdotz1  [1] rdotz2  [2] alpha  [3] beta  [4] pAp  [5] rdotr  [6] zdotAp  [7] TOL
 *  [8] STOP  [9] ITERS.
 * libp_pcg_update_f64 counts ITERS and sets STOP once rdotr <= TOL (TOL > 0
 * only); every fused PCG kernel returns at once while STOP is set. Iterations
 * can so be queued in chunks and the host read the scalars once per chunk,
 * with the same arithmetic as a convergence test after every iteration. */
enum { LIBP_PCG_RDOTZ1 = 0, LIBP_PCG_RDOTZ2, LIBP_PCG_ALPHA, LIBP_PCG_BETA, LIBP_PCG_PAP,
       LIBP_PCG_RDOTR, LIBP_PCG_ZDOTAP, LIBP_PCG_TOL, LIBP_PCG_STOP, LIBP_PCG_ITERS, LIBP_PCG_NSCALARS };

/* Replaces updatePCG (libs/linearSolver/okl/linearSolverUpdatePCG.okl:29-75) and the
 * host sum of pcg::UpdatePCG (linearSolverPCG.cpp:153-171): alpha = s[RDOTZ1]/s[PAP]
 * is formed on the device, then x += alpha*p, r -= alpha*Ap, s[RDOTR] = r.r. */
int libp_pcg_update_f64(int N, const double* p, const double* Ap, double* x, double* r, double* scalars,
                        void* ws, void* stream);
/* r.z and (flexible) z.Ap, then beta and p = z + beta*p, as in
 * linearSolverPCG.cpp:113-127 (iter==0 gives beta=0). */
int libp_pcg_rdotz_f64(int N, const double* r, const double* z, const double* Ap, int flexible, double* scalars,
                       void* ws, void* stream);
/* The same with z given in FP32 (the output of an FP32 preconditioner): z is
 * widened into the FP64 vector in the pass that forms the scalars; results
 * are those of libp_convert_f32_to_f64 followed by libp_pcg_rdotz_f64. */
int libp_pcg_rdotz_from_f32_f64(int N, const double* r, const float* z32, double* z, const double* Ap, int flexible,
                                double* scalars, void* ws, void* stream);
int libp_pcg_update_p_f64(int N, int first_iter, const double* z, double* p, const double* scalars, void* stream);
/* p.Ap into scalars[PAP]. */
int libp_pcg_pap_f64(int N, const double* p, const double* Ap, double* scalars, void* ws, void* stream);
/* The gather of the operator fused with p.Ap (ellipticOperator.cpp:29-107's
 * gather followed by linearSolverPCG.cpp:134-140): gv = gather(v) over the
 * CSR rows, bit-identical to libp_ogs_gather_f64, and scalars[PAP] = p.gv,
 * scalars[ALPHA] = scalars[RDOTZ1] / p.gv. */
int libp_pcg_gather_pap_f64(int Nrows, const int* rowStarts, const int* colIds, const double* v, double* gv,
                            const double* p, double* scalars, void* ws, void* stream);
/* The update of libp_pcg_update_f64 fused with the Jacobi preconditioner and
 * the r.z of the next iteration (linearSolverPCG.cpp:107-124,142-160 with
 * ellipticPreconJacobi.cpp:42-51): x += alpha p, r -= alpha Ap,
 * z = invDiag.*r, scalars[RDOTR] = r.r, ITERS and STOP as in
 * libp_pcg_update_f64, then RDOTZ2 <- RDOTZ1 <- r.z and BETA = r.z / RDOTZ2. */
int libp_pcg_jacobi_update_f64(int N, const double* invDiag, const double* p, const double* Ap, double* x, double* r,
                               double* z, double* scalars, void* ws, void* stream);
/* Allocation-free variant of libp_pcg_jacobi_update_f64 (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code under
 * HIP graph capture, tight solve loops and harnesses that forbid hipMalloc. The
 * block partials go to a 32 KiB static device array (one per GPU, held until the
 * process ends) instead of a caller workspace; nothing is allocated. Results are
 * bit-identical to libp_pcg_jacobi_update_f64. The partial count is capped at 1024
 * blocks, so every N in [0, LIBP_NOALLOC_MAX_N] fits; other N are rejected with
 * LIBP_ERR_ARG. All calls on one GPU share that GPU's array: calls that overlap
 * on one GPU (two streams, or two threads) race and must be serialised by the
 * caller. The stream must belong to the current device. */
#define LIBP_NOALLOC_MAX_N 2147483647L
long libp_pcg_jacobi_update_noalloc_max_n(void);
int libp_pcg_jacobi_update_noalloc_f64(int N, const double* invDiag, const double* p, const double* Ap, double* x,
                                       double* r, double* z, double* scalars, void* stream);
/* A Jacobi-PCG iteration on the C0 hex operator in three passes: the p update
 * (libp_pcg_update_p_f64), the Ax with p.Ap, and the gather fused with the
 * update (no stored Ap). Replaces, for PCG with the Jacobi preconditioner,
 * the Operator call of linearSolverPCG.cpp:130 (ellipticOperator.cpp:29-107:
 * partialAxKernel, then the ogs gather), the p.Ap dot at :134-140 and updatePCG
 * with the preconditioner and r.z at :142-160 (ellipticPreconJacobi.cpp:42-51).
 *
 * libp_pcg_ax_hex_pap_f64: the partial Ax of ellipticAxHex3D.okl:156-295
 * applied to the gathered p (through GlobalToLocal) into the E-vector AqL,
 * then scalars[PAP] = p.Ap summed from the local nodes (sum of p_L * AqL,
 * which equals p.Ap in exact arithmetic) and scalars[ALPHA] =
 * scalars[RDOTZ1] / p.Ap. papPartials holds libp_ax_hex_split_blocks(N,
 * Nelements) doubles (Nelements at N=15, where the FP64 MFMA kernel of
 * libp_ax_hex_partial_mfma_n15 runs, one element per workgroup). N in [1,15].
 *
 * libp_pcg_gather_jacobi_update_f64: per gathered row, Ap = the sum of the
 * row's copies of AqL (the gather's arithmetic), then x += alpha p,
 * r -= alpha Ap, z = invDiag.*r, r.r and r.z; ITERS, STOP, RDOTZ1/RDOTZ2 and
 * BETA as in libp_pcg_jacobi_update_f64. partials holds
 * libp_pcg_gather_jacobi_partials(Nrows) doubles. Both return at once while
 * scalars[STOP] is set. */
int libp_pcg_ax_hex_pap_f64(int N, int Nelements, const int* elementList, const int* GlobalToLocal, const double* wJ,
                            const double* ggeo, const double* D, double lambda, const double* p, double* AqL,
                            double* papPartials, double* scalars, void* stream);
long libp_pcg_gather_jacobi_partials(int Nrows);
int libp_pcg_gather_jacobi_update_f64(int Nrows, const int* rowStarts, const int* colIds, const double* AqL,
                                      const double* invDiag, const double* p, double* x, double* r, double* z,
                                      double* scalars, double* partials, void* stream);
/* scalars[TOL] = tol^2 * scalars[RDOTR] (1e-300 when that is 0), and
 * scalars[STOP] = 1 when scalars[RDOTR] is 0: the pure relative stopping rule
 * of a solve from a zero guess, formed on the device instead of after a host
 * read of r0.r0 (linearSolverPCG.cpp:93-97, 107-108). */
int libp_pcg_relative_tol_f64(double tol, double* scalars, void* stream);
/* Jacobi preconditioner z = invDiag.*r fused with r.z (JacobiPrecon::Operator,
 * solvers/elliptic/src/ellipticPreconJacobi.cpp:42-51 followed by the r.z dot). */
int libp_pcg_jacobi_rdotz_f64(int N, const double* invDiag, const double* r, double* z, const double* Ap,
                              int flexible, double* scalars, void* ws, void* stream);

/* One Jacobi-PCG iteration on the N=1 hex operator in three launches (no
 * single reference counterpart: it fuses the iteration of
 * linearSolverPCG.cpp:107-171 with JacobiPrecon::Operator,
 * ellipticPreconJacobi.cpp:42-51, and the N=1 partial Ax + gather of
 * ellipticOperator.cpp:29-107). Used for the p-MG coarse level. Before the
 * first call: scalars zeroed, scalars[TOL] set, z = invDiag.*r and
 * scalars[RDOTZ1] = r.z (libp_pcg_jacobi_rdotz_f64). first = 1 on the first
 * iteration. AqL holds Nelements*8 doubles. Each reduction is finished by
 * the next launch, so the stopping test of an iteration runs at the start of
 * the next one (then every launch returns at once via scalars[STOP]); call
 * libp_pcg_jacobi_hex1_check_f64 after the last queued iteration before
 * reading scalars[STOP], scalars[RDOTR] or scalars[ITERS]. ws must not be
 * used by other reductions between the calls. */
int libp_pcg_jacobi_hex1_iteration_f64(int Nelements, const int* GlobalToLocal, const double* wJ, const double* ggeo,
                                       const double* D, double lambda, int Nrows, const int* rowStarts,
                                       const int* colIds, const double* invDiag, int first, double* x, double* r,
                                       double* z, double* p, double* Ap, double* AqL, double* scalars, void* ws,
                                       void* stream);
/* Stopping test of the last libp_pcg_jacobi_hex1_iteration_f64: sets
 * scalars[RDOTR] = r.r and, when r.r <= scalars[TOL], scalars[STOP] = 1. */
int libp_pcg_jacobi_hex1_check_f64(int Nrows, double* scalars, void* ws, void* stream);
/* The same Jacobi-PCG on the N=1 hex operator in its single-reduction form
 * (Chronopoulos and Gear 1989; the PCG iterates of linearSolverPCG.cpp:107-171
 * in exact arithmetic, with ellipticPreconJacobi.cpp:42-51 and the N=1
 * ellipticAxHex3D.okl:156-295): gamma = r.u and delta = w.u (u = invDiag.*r,
 * w = A u) in one reduction, two launches per iteration. The vector updates
 * run inside the next Ax launch; owner (one byte per local node, 1 at the
 * first local copy of each gathered node) says which copy writes them. p2,
 * s2, r2, u2: host arrays of two device vectors each, used by iteration
 * parity. mode 0 starts from x = 0 with the right-hand side in r2[0] (parity
 * 0); mode 1 is iteration 0 (parity 0); mode 2 iteration i >= 1 with parity
 * i % 2; mode 3 applies the stopping test (r.r <= tol^2 r0.r0) of the last
 * queued iteration before the host reads scalars[STOP], [ITERS], [RDOTR].
 * scalars zeroed before mode 0; ws must not be used by other reductions
 * between the calls. */
int libp_pcg_jacobi_hex1_cg1_f64(int Nelements, const int* GlobalToLocal, const unsigned char* owner, const double* wJ,
                                 const double* ggeo, const double* D, double lambda, int Nrows, const int* rowStarts,
                                 const int* colIds, const double* invDiag, double tol, int mode, int parity, double* x,
                                 double* w, double* const* p2, double* const* s2, double* const* r2,
                                 double* const* u2, double* AqL, double* scalars, void* ws, void* stream);

/* The iteration of libp_pcg_jacobi_hex1_cg1_f64 on the assembled operator
 * (CSR over the gathered dofs, columns ascending; the matrix of the N=1
 * level as ellipticBuildOperatorMatrixContinuous assembles it for parAlmond,
 * ellipticPreconMultiGrid.cpp:104-131): one launch per iteration, each row
 * recomputing the updated u at its columns. Modes, parity, scalars and ws
 * as there, except: w2 holds two device vectors used by parity, as p2; ws
 * holds two reduction workspaces (2 * libp_reduce_ws_bytes()), alternated
 * by parity; mode 3 takes the parity of the last queued iteration. Same
 * iterates in exact arithmetic; in floating point A u is summed by rows of
 * the assembled matrix instead of by elements. */
int libp_pcg_jacobi_csr_cg1_f64(int Nrows, const int* rowStarts, const int* colIds, const double* vals,
                                const double* invDiag, double tol, int mode, int parity, double* x, double* const* w2,
                                double* const* p2, double* const* s2, double* const* r2, double* const* u2,
                                double* scalars, void* ws, void* stream);
/* The stopping test of mode 3 of libp_pcg_jacobi_csr_cg1_f64 (parity of the
 * last queued iteration), which then copies scalars[0, LIBP_PCG_NSCALARS) into
 * hostMirror, pinned host memory mapped for the device (its device pointer),
 * so that a host check after a chunk of iterations needs no further launch. */
int libp_pcg_jacobi_csr_cg1_check_f64(int Nrows, int parity, double* scalars, void* ws, double* hostMirror,
                                      void* stream);

/* Non-blocking solvers (libs/linearSolver/linearSolverNBPCG.cpp,
 * linearSolverNBFPCG.cpp): fused update + multi-dot kernels replacing
 * update1NBPCG/update2NBPCG (okl/linearSolverUpdateNBPCG.okl) and
 * update0NBFPCG/update1NBFPCG (okl/linearSolverUpdateNBFPCG.okl) together
 * with the host sums and scalar recurrences that follow them. The scalars stay
 * on the device:
 *   NBPCG  [0] gamma=r.z [1] previous gamma [2] alpha [3] beta [4] delta=p.s [5] r.r [6] z.z
 *   NBFPCG [0] gamma=u.r [1] delta=u.w [2] eta [3] alpha [4] beta [5] r.r
 * nbpcg_update1: p = z + beta p, s = Z + beta s, delta = p.s, alpha = gamma/delta.
 * nbpcg_update2: r -= alpha s, z -= alpha S, r.z/z.z/r.r, beta = gamma/previous gamma
 *   (first = 1: alpha = beta = 0, linearSolverNBPCG.cpp:86-90).
 * nbfpcg_update0: u.r, u.w, r.r; eta = delta; alpha = gamma/eta.
 * nbfpcg_update1: x += alpha p, r -= alpha s, u -= alpha q, w -= alpha z; u.r, u.s,
 *   u.w, r.r; beta = -u.s/eta; eta = delta - beta^2 eta; alpha = gamma/eta.
 * nbfpcg_update2: p = u + beta p, s = w + beta s, q = m + beta q, z = n + beta z
 *   (linearSolverNBFPCG.cpp:155-165). axpy_dev: y += (*alpha) x. */
int libp_nbpcg_update1_f64(int N, const double* z, const double* Z, double* p, double* s, double* scalars, void* ws,
                           void* stream);
int libp_nbpcg_update2_f64(int N, int first, const double* s, const double* S, double* r, double* z, double* scalars,
                           void* ws, void* stream);
int libp_axpy_dev_f64(int N, const double* alpha, const double* x, double* y, void* stream);
int libp_nbfpcg_update0_f64(int N, const double* u, const double* r, const double* w, double* scalars, void* ws,
                            void* stream);
int libp_nbfpcg_update1_f64(int N, const double* p, const double* s, const double* q, const double* z, double* x,
                            double* r, double* u, double* w, double* scalars, void* ws, void* stream);
int libp_nbfpcg_update2_f64(int N, const double* u, const double* w, const double* m, const double* n, double* p,
                            double* s, double* q, double* z, const double* scalars, void* stream);

/* ------------------------------------------------------------------ */
/* p-multigrid transfers and smoother pieces                            */
/* ------------------------------------------------------------------ */
/* Replaces ellipticPartialPreconCoarsenHex3D (okl/ellipticPreconCoarsenHex3D.okl:209-294),
 * called by MGLevel::coarsen (src/ellipticPreconMultiGridLevel.cpp:41-82):
 * qc[e*NpC+..] = (P^T x P^T x P^T) qf_L, qf_L gathered through the fine GlobalToLocal.
 * P is NqF x NqC row-major (P[f*NqC+c]). */
int libp_mg_coarsen_hex_f64(int Nf, int Nc, int Nelements, const int* elementList, const int* GlobalToLocal,
                            const double* P, const double* qf, double* qc, void* stream);
int libp_mg_coarsen_hex_f32(int Nf, int Nc, int Nelements, const int* elementList, const int* GlobalToLocal,
                            const float* P, const float* qf, float* qc, void* stream);
/* The coarsening with the fine residual's inverse-multiplicity weight
 * (ellipticPreconMultiGridLevel.cpp:51) applied as the kernel loads it:
 * coarsen(weight .* qf), bit-identical to an amxpy pass followed by the
 * coarsening. */
int libp_mg_coarsen_weighted_hex_f64(int Nf, int Nc, int Nelements, const int* elementList, const int* GlobalToLocal,
                                     const double* P, const double* weight, const double* qf, double* qc,
                                     void* stream);
int libp_mg_coarsen_weighted_hex_f32(int Nf, int Nc, int Nelements, const int* elementList, const int* GlobalToLocal,
                                     const float* P, const float* weight, const float* qf, float* qc, void* stream);
/* Replaces ellipticPartialPreconProlongateHex3D (okl/ellipticPreconProlongateHex3D.okl:216-301),
 * called by MGLevel::prolongate (ellipticPreconMultiGridLevel.cpp:84-125). */
int libp_mg_prolong_hex_f64(int Nf, int Nc, int Nelements, const int* elementList, const int* GlobalToLocalC,
                            const double* P, const double* qc, double* qf, void* stream);
int libp_mg_prolong_hex_f32(int Nf, int Nc, int Nelements, const int* elementList, const int* GlobalToLocalC,
                            const float* P, const float* qc, float* qf, void* stream);
/* Chebyshev recurrence steps of MGLevel::smoothChebyshev (ellipticPreconMultiGridLevel.cpp:154-206):
 * cheb_start: RES = invDiag.*(r - Ax) (Ax may be NULL: x is zero), d = invTheta*RES, and, when
 *             xIsZero, X = d;
 * cheb_step:  RES -= invDiag.*Ad, d = c1*RES + c2*d, X += d (fused).
 * Both FP32 (the smoother precision) and FP64. */
int libp_cheb_start_f32(int N, const float* invDiag, const float* r, const float* Ax, float invTheta, float* RES,
                        float* d, float* X, int xIsZero, void* stream);
int libp_cheb_step_f32(int N, const float* invDiag, const float* Ad, float c1, float c2, float* RES, float* d,
                       float* X, void* stream);
int libp_cheb_start_f64(int N, const double* invDiag, const double* r, const double* Ax, double invTheta,
                        double* RES, double* d, double* X, int xIsZero, void* stream);
int libp_cheb_step_f64(int N, const double* invDiag, const double* Ad, double c1, double c2, double* RES,
                       double* d, double* X, void* stream);

/* Algebraic multigrid of the p-MG base level (libs/parAlmond):
 * csr_spmv: y = alpha A x + beta z for a CSR matrix (z may equal y, or be NULL with
 *           beta = 0); lanesPerRow 1, 8 or 64 lanes reduce each row in a fixed order
 *           (replaces SpMVcsr, okl/SpMVcsr.okl, and the Chebyshev CSR kernels of
 *           parAlmondAMGSmoother.cpp:68-160);
 * dense_gemv: y = M x with M dense row-major n x n, the coarsest level's inverse
 *           (parAlmondCoarseExact.cpp:35-78, okl/dGEMV.okl). */
int libp_csr_spmv_f64(int Nrows, const int* rowStarts, const int* cols, const double* vals, double alpha,
                      const double* x, double beta, const double* z, double* y, int lanesPerRow, void* stream);
int libp_dense_gemv_f64(int n, const double* M, const double* x, double* y, void* stream);

/* ------------------------------------------------------------------ */
/* Setup kernels (mesh geometry, operator diagonal, right-hand side)    */
/* ------------------------------------------------------------------ */
/* Physical nodes and volume geometric factors of hex elements: restates
 * meshPhysicalNodesHex3D.cpp:31-112 (trilinear map of EX/EY/EZ [e][8]) followed by
 * the optional curved map x += amp*sin(2 pi y)*sin(2 pi z) (no reference
 * counterpart), and meshGeometricFactorsHex3D.cpp:31-193 for ggeo and wJ.
 * r1 = GLL nodes (Nq), w1 = GLL weights (Nq), D as above. *badElement receives
 * 1 + the largest element index whose Jacobian fell below 1e-12 (0 if none);
 * it must be zero on entry (device int). */
int libp_geofacs_hex_f64(int N, int Nelements, const double* EX, const double* EY, const double* EZ,
                         const double* r1, const double* w1, const double* D, double curveAmp, double* x,
                         double* y, double* z, double* ggeo, double* wJ, int* badElement, void* stream);
/* Local (E-vector) operator diagonal of BuildOperatorDiagonalContinuousHex3D
 * (src/ellipticBuildOperatorDiagonal.cpp:998-1057); mapB is the elliptic node flag
 * (1 = Dirichlet). The caller gathers it with libp_ogs_gather_f64 (Trans). */
int libp_diag_hex_local_f64(int N, int Nelements, const double* ggeo, const double* wJ, const double* D,
                            double lambda, const int* mapB, int allNeumann, double neumannShift, double* diagL,
                            void* stream);
/* Forcing of ellipticRhsHex3D.okl:28-51 for the built-in data sets, and the Dirichlet
 * data used by ellipticRhsBCHex3D / ellipticAddBCHex3D:
 *   dataKind 0: data/ellipticSine3D.h (f=(3pi^2+lambda) sin sin sin, u_b = sin sin sin)
 *   dataKind 1: seeded 64-mode Fourier forcing (coeffs: 64 x {kx,ky,kz,amp,phase}), u_b = 0.
 * rhsL = wJ*f; ubL = (mapB==1) ? u_b : 0. */
int libp_rhs_hex_f64(int Nelements, int Np, int dataKind, const double* coeffs, const double* x, const double* y,
                     const double* z, const double* wJ, double lambda, const int* mapB, double* rhsL, double* ubL,
                     void* stream);
/* xL[n] = u_b(x,y,z) where mapB==1 (ellipticAddBCHex3D.okl:27-47). */
int libp_addbc_hex_f64(int NL, int dataKind, const double* coeffs, const double* x, const double* y, const double* z,
                       const int* mapB, double* xL, void* stream);

/* ------------------------------------------------------------------ */
/* Learned per-element smoother (no reference counterpart)              */
/* ------------------------------------------------------------------ */
/* None of these entry points has a counterpart in the reference: the
 * reference has no learned components. The network is the per-element MLP
 * nIn = Np + 6*Nfp -> 256 -> 256 -> Np with ReLU, BF16 weights and FP32
 * accumulation on v_mfma_f32_32x32x16_bf16. Matrices are row-major with the
 * given leading dimensions (in elements); BF16 values are passed as
 * unsigned short bit patterns. */

/* One layer forward: Y[M x Nout] = act(X[M x K] Wt^T + b), X BF16 [M][ldx],
 * Wt BF16 [Nout][ldw] (PyTorch's Linear.weight layout), b FP32 [Nout] or NULL.
 * relu = 1: Y is BF16 = bf16(max(., 0)); relu = 0: Y is FP32 (no activation). */
int libp_mlp_linear_fwd(int M, int K, int Nout, const unsigned short* X, long ldx, const unsigned short* Wt,
                        long ldw, const float* b, int relu, void* Y, long ldy, void* stream);
/* The FP32 layer of libp_mlp_linear_fwd (relu = 0) with each output row
 * multiplied by rowScale[m] in the epilogue: Y[m][n] = rowScale[m] *
 * fl((X Wt^T + b)[m][n]), the same product that the assembly forms from an
 * unscaled Y, so libp_learned_assemble_f32 with scale NULL on this Y gives the
 * same result bit for bit with one load per copy instead of two. */
int libp_mlp_linear_fwd_scaled(int M, int K, int Nout, const unsigned short* X, long ldx, const unsigned short* Wt,
                               long ldw, const float* b, const float* rowScale, float* Y, long ldy, void* stream);
/* Backward through the weights: dX[M x K] = bf16((dY Wt) .* (H > 0)), dY BF16
 * [M][lddy], H BF16 [M][ldh] (the layer input after its ReLU) or NULL for no mask. */
int libp_mlp_linear_bwd_data(int M, int K, int Nout, const unsigned short* dY, long lddy, const unsigned short* Wt,
                             long ldw, const unsigned short* H, long ldh, unsigned short* dX, long lddx,
                             void* stream);
/* Weight gradients: dWt[Nout x K] = dY^T X (FP32, [Nout][K] dense) and
 * db[Nout] = column sums of dY (FP32, or NULL). Split-K over M with FP32
 * partials summed in a fixed order; ws holds libp_mlp_ws_bytes(M, K, Nout). */
int libp_mlp_linear_bwd_weight(int M, int K, int Nout, const unsigned short* dY, long lddy, const unsigned short* X,
                               long ldx, float* dWt, float* db, void* ws, void* stream);
long libp_mlp_ws_bytes(int M, int K, int Nout);
/* Adam (Kingma & Ba 2015) on FP32 masters, writing a BF16 copy (wbf16 may be
 * NULL); step counts from 1. */
int libp_adam_step(int n, float* master, unsigned short* wbf16, const float* grad, float* m, float* v, float lr,
                   float beta1, float beta2, float eps, int step, void* stream);
int libp_f32_to_bf16(long n, const float* x, unsigned short* y, void* stream);

/* Network input: X[e][0:Np] = z at the element's nodes, X[e][Np:Np+Ntr] = z at
 * traceIds[e][:] (gathered ids, -1 = none), z = invDiag .* r, each row divided
 * by its RMS, which is written to scale[e] (0 for an all-zero row). */
int libp_learned_features_f32(int Nelements, int Np, int Ntr, const int* GlobalToLocal, const int* traceIds,
                              const float* r, const float* invDiag, unsigned short* X, long ldx, float* scale,
                              void* stream);
/* Weights W[Nout][ldw] (BF16, K inputs) in the order the fused layer reads them:
 * 16-byte chunk (cb*K/16 + t)*64 + lane = W[32 cb + lane%32][16 t + 8 (lane/32) .. +7].
 * Nout a multiple of 32, K of 16; Wpacked holds Nout*K values, 16-byte aligned. */
int libp_learned_pack_weights_bf16(int Nout, int K, const unsigned short* W, long ldw, unsigned short* Wpacked,
                                   void* stream);
/* The linear smoother's layer fused with its input: Y[e][n] = scale[e] * sum_k X[e][k] W[n][k]
 * with X and scale as libp_learned_features_f32 forms them (scale may be NULL), W packed
 * by libp_learned_pack_weights_bf16. The same Y, to the bit, as libp_learned_features_f32
 * followed by libp_mlp_linear_fwd_scaled (bias NULL); X is never stored. N=7 shape only:
 * Np = 512, Ntr = 384, Nout = 512 (other shapes return an error). */
int libp_learned_linear_layer_f32(int Nelements, int Np, int Ntr, int Nout, const int* GlobalToLocal,
                                  const int* traceIds, const float* r, const float* invDiag,
                                  const unsigned short* Wpacked, float* Y, long ldy, float* scale, void* stream);
/* c[g] = (1/|row g|) sum_{j in row g} scale[j/Np] Y[j/Np][j%Np] + omega invDiag[g] r[g],
 * rows of the masked gather (rowStarts, colIds); omega is a device pointer (or NULL = 0).
 * scale NULL: Y holds the scaled values already (libp_mlp_linear_fwd_scaled). */
int libp_learned_assemble_f32(int Nrows, const int* rowStarts, const int* colIds, int Np, const float* Y, long ldy,
                              const float* scale, const float* omega, const float* r, const float* invDiag, float* c,
                              void* stream);
/* The same assembly added to c (c += correction: the smoother's x += c). */
int libp_learned_assemble_add_f32(int Nrows, const int* rowStarts, const int* colIds, int Np, const float* Y,
                                  long ldy, const float* scale, const float* omega, const float* r,
                                  const float* invDiag, float* c, void* stream);
/* Energy-optimal step: sc[0] = c.res, sc[1] = c.Ac, sc[2] = alpha = sc[0]/sc[1]
 * (0, and sc[3] += 1, if not finite or c.Ac <= 0). ws: libp_reduce_ws_bytes(). */
int libp_learned_alpha_f32(int N, const float* c, const float* res, const float* Ac, double* sc, void* ws,
                           void* stream);
/* x = (xIsZero ? 0 : x) + sc[2] c; res -= sc[2] Ac when res != NULL. */
int libp_learned_update_f32(int N, int xIsZero, const double* sc, const float* c, const float* Ac, float* x,
                            float* res, void* stream);
/* Training loss J = d.Ad / *nrm of one sample (d = target - c, Ad = A d) and
 * gc = dJ/dc = -2 Ad / *nrm; stat[0] = J, stat[1] += J; *gradOmega = gc.(invDiag.*r). */
int libp_learned_loss_f32(int N, const float* d, const float* Ad, const float* r, const float* invDiag,
                          const double* nrm, float* gc, double* stat, float* gradOmega, void* ws, void* stream);
/* dY[e][n] = bf16(gc[g] scale[e] invLen[g]) at g = GlobalToLocal[e*Np+n] >= 0, else 0. */
int libp_learned_scatter_grad_f32(int Nelements, int Np, const int* GlobalToLocal, const float* gc,
                                  const float* invLen, const float* scale, unsigned short* dY, long lddy,
                                  void* stream);
/* Linear smoother training target T[e][n] = bf16(eh[g] / scale[e]), g = GlobalToLocal[e*Np+n]
 * (0 where g < 0). */
int libp_learned_element_target_f32(int Nelements, int Np, const int* GlobalToLocal, const float* eh,
                                    const float* scale, unsigned short* T, long ldt, void* stream);
/* x[i] = uniform(-1, 1), a pure function of (seed, i) (splitmix64). */
int libp_random_uniform_f32(long N, unsigned long long seed, float* x, void* stream);
/* invLen[g] = 1 / (rowStarts[g+1] - rowStarts[g]). */
int libp_inv_row_length_f32(int Nrows, const int* rowStarts, float* invLen, void* stream);

/* ------------------------------------------------------------------ */
/* Solver handle: the C++ elliptic_t driven through the C ABI           */
/* ------------------------------------------------------------------ */
/* Creates platform + mesh + elliptic_t (+ preconditioner) from .rc text in the
 * reference format (libs/core/settings.cpp:134-214). Returns NULL on error. */
void* libp_elliptic_create(const char* rc_text);
void libp_elliptic_free(void* h);
/* Scalar facts about the handle: "Nelements","N","Np","Ngather","Nmasked",... */
long libp_elliptic_geti(void* h, const char* name);
double libp_elliptic_getd(void* h, const char* name);
/* Device pointer of a named array of the fine level ("ggeo","wJ","GlobalToLocal",
 * "rowStartsT","colIdsT","rowStartsN","colIdsN","x","y","z","D","invDiag",...),
 * with its length and dtype (0 f64, 1 i32, 2 i64, 3 f32). Host arrays from setup
 * ("globalIds","mapB","blockRowStartsT","blockRowStartsN","EToV") are returned
 * as host pointers with *on_host = 1. */
void* libp_elliptic_array(void* h, const char* name, long* length, int* dtype, int* on_host);
/* The HIP stream the handle launches on. */
void* libp_elliptic_stream(void* h);
/* Aq = A q on gathered vectors (elliptic_t::Operator). */
int libp_elliptic_operator(void* h, const double* q, double* Aq);
/* Partial Ax only (E-vector out). */
int libp_elliptic_partial_ax(void* h, const double* q, double* AqL);
/* Apply the configured preconditioner: z = M^{-1} r. */
int libp_elliptic_precon(void* h, const double* r, double* z);
/* elliptic_t::Run: rhs, solve, boundary fill, solution norm. Results in out[]:
 * [0] solution norm, [1] iterations, [2] solve seconds, [3] final |r|, [4] |r0|,
 * [5] max nodal error vs the manufactured solution (sine data only, else NaN). */
int libp_elliptic_run(void* h, double* out, int nout);
/* Copy of the solution E-vector (host buffer of Nelements*Np doubles) from the last run. */
int libp_elliptic_solution(void* h, double* xL_host);

/* Learned fine smoother of a handle built with MULTIGRID SMOOTHER LEARNED and
 * MULTIGRID PRECISION FP32 (no reference counterpart).
 * Training: options "key=value" separated by blanks (seeds=101,102
 * daggerSeeds=201 steps=3000 dagger=1 daggerSteps=1500 lr=1e-3 targetTol=1e-9
 * initSeed=1 logEvery=250 maxRecord=24 verbose=1; for MULTIGRID LEARNED ARCH
 * LINEAR: samples=16 ridge=1e-4 coarseTol=1e-6 initSeed=1 verbose=1). Returns
 * the number of loss values (MLP: mean loss per logEvery steps; LINEAR: the
 * validation loss of the fit; up to nloss copied), or -status. */
int libp_elliptic_learned_train(void* h, const char* options, double* loss, int nloss);
/* Weight files: header + FP32 parameters + FNV-1a hash; a rejected file leaves
 * the Chebyshev smoother in use and returns LIBP_ERR_SETUP (reason in
 * libp_last_error and libp_elliptic_learned_status). */
int libp_elliptic_learned_load(void* h, const char* path);
int libp_elliptic_learned_save(void* h, const char* path);
/* out: [0] active, [1] fallen back to Chebyshev, [2] rejected steps, [3] parameter
 * count, [4] last step length alpha, [5] last training loss, [6] input width, [7] hidden width,
 * [8] architecture (0 MLP, 1 LINEAR), [9] sweeps */
int libp_elliptic_learned_info(void* h, double* out, int nout);
const char* libp_elliptic_learned_status(void* h);
/* FP32 parameter vector (host buffers of the parameter count). */
int libp_elliptic_learned_get_params(void* h, float* theta, long n);
int libp_elliptic_learned_set_params(void* h, const float* theta, long n, int activate);
/* Network correction c(r) of a device FP32 gathered residual (before the step length). */
int libp_elliptic_learned_correction(void* h, const float* r, float* c);
/* Training loss of one sample (device FP32 r, target eh; nrm = eh.A eh) and its
 * gradient, left in the handle's "learnedGrad" device array; no update. */
int libp_elliptic_learned_loss_grad(void* h, const float* r, const float* eh, double nrm, double* loss);
/* Assembled operator of the handle's fine level (N <= 3) on the gathered dofs, CSR:
 * libp_elliptic_csr_nnz returns the nonzero count (or -status); libp_elliptic_csr fills
 * host arrays rowStarts [Ngather+1], cols [nnz], vals [nnz]
 * (ellipticBuildOperatorMatrixContinuous restated through the element operator). */
long libp_elliptic_csr_nnz(void* h);
int libp_elliptic_csr(void* h, int* rowStarts, int* cols, double* vals);
/* Smoothed-aggregation AMG on the handle's fine operator (N <= 3, not all-Neumann):
 * setup, then x = V(b) for device FP64 gathered vectors (one V-cycle from zero);
 * info: [0] levels, then rows and nonzeros of each level (up to nout values). */
int libp_elliptic_amg_setup(void* h);
int libp_elliptic_amg_vcycle(void* h, const double* b, double* x);
int libp_elliptic_amg_info(void* h, double* out, int nout);
/* Host copy of a named array of AMG level `level`, for tests that compare the setup
 * with the CPU restatement (parAlmondAMGSetup.cpp:35-144). which: 0 the AMG of
 * libp_elliptic_amg_setup, 1 the base level of the handle's p-MG ([MULTIGRID COARSE
 * SOLVER] AMG). Names: A_rs, A_ci, P_rs, P_ci, R_rs, R_ci, aggregate (int32);
 * A_v, P_v, R_v, invDiag, null, rho = {rho, lambda0, lambda1, lambda} (float64).
 * out may be null to query *length; *dtype is 0 float64 or 1 int32. */
int libp_elliptic_amg_level_array(void* h, int which, int level, const char* name, void* out, long* length,
                                  int* dtype);
/* One smoothing of level `level` of the libp_elliptic_amg_setup AMG
 * (amgLevel::smooth, parAlmondAMGLevel.cpp:63-72): device b and x (updated in place). */
int libp_elliptic_amg_smooth(void* h, int level, const double* b, double* x, int x_is_zero);
/* p-MG level access, for tests that compare each level operation with the CPU
 * restatement. ndofs returns level `level`'s gathered length (or -status) and the
 * level precision in *bits (32 or 64). op: 0 out = A in (MGLevel::Operator,
 * ellipticPreconMultiGridLevel.cpp:30-32); 1 out = R in (coarsen, :41-82); 2 out +=
 * P in (prolongate, :84-125); 3 smoother (:127-206; in = rhs, out = x updated in
 * place, flag = x is zero); 4 out = inverse diagonal. Device vectors of the level
 * precision. lambda returns the Chebyshev interval {lambda1, lambda0} (:361-368). */
long libp_elliptic_mg_level_ndofs(void* h, int level, int* bits);
int libp_elliptic_mg_level_op(void* h, int level, int op, const void* in, void* out, int flag);
int libp_elliptic_mg_level_lambda(void* h, int level, double* lambda);
/* LINEAR architecture: Gram matrices of one least-squares sample (device FP32 r
 * and eh; host G [nIn][nIn] = X^T X, B [Np][nIn] = T^T X with X the BF16 network
 * input of r and T[e][n] = bf16(eh / s_e) at the element's nodes). */
int libp_elliptic_learned_linear_gram(void* h, const float* r, const float* eh, double* G, double* B);

#ifdef __cplusplus
}
#endif

#endif
