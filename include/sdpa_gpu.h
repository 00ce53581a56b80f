/*
 * sdpa_gpu.h -- C ABI of the MI355X (gfx950) Newton-direction hot path of
 * SDPA 7.4.4.  This is the drop-in boundary: the host interior-point loop
 * (SDPA::solve, sdpa_solve.cpp:59-233) calls only these entry points, at the
 * seams where the reference calls Newton::Mehrotra, StepLength::
 * computeStepLength and Solutions::update (sdpa_solve.cpp:96-197).
 *
 * Conventions
 *  - Plain C types only; every entry returns int: 0 = success, > 0 = a
 *    LAPACK-style info (the 1-based failing pivot, or 1 for "the X/Z
 *    factorization failed"), < 0 = an error (see sdpa_gpu_last_error).
 *    Nothing here calls exit().
 *  - Problem data follow the .dat-s records (sdpa_io.cpp:473-606): k = 0..m
 *    (0 is F0), l = 1..nBlock, i, j 1-based, LP blocks have negative size.
 *    Internally C = -F0, A_k = F_k, b = c (the reference's reversed
 *    primal/dual, sdpa_tool.h:95).
 *  - Matrices in the "packed block space" are the SDP blocks column-major,
 *    concatenated in block order, followed by the LP coordinates.  All
 *    iterate arrays (X, Z, Dx, Dz, ...) stay resident on the device; host
 *    copies move only through set/get calls.
 *  - One context per host thread; it owns one HIP stream and all buffers.
 */
#ifndef SDPA_GPU_H
#define SDPA_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdpa_gpu_ctx sdpa_gpu_ctx;

/* Context lifetime.  No reference counterpart (the reference keeps this state
 * inside class SDPA, sdpa_call.h:91-267). */
int sdpa_gpu_create(int device, sdpa_gpu_ctx** ctx);
int sdpa_gpu_destroy(sdpa_gpu_ctx* ctx);
const char* sdpa_gpu_last_error(sdpa_gpu_ctx* ctx);

/* Load a problem.  Replaces IO::read/SDPA::inputElement + initializeUpperTriangle
 * (sdpa_io.cpp:253-606, sdpa_call.cpp:285-341, 650-666) and the structural part
 * of SDPA::initializeSolve (sdpa_solve.cpp:10-29): changeToDense
 * (sdpa_struct.cpp:510-555), InputData::initialize_index (sdpa_dataset.cpp:268-400),
 * Chordal::ordering_bMat's dense decision (sdpa_chordal.cpp:365-410) and
 * Newton::computeFormula_SDP with kappa (sdpa_newton.cpp:515-695). */
int sdpa_gpu_load_problem(sdpa_gpu_ctx* ctx, int m, int nBlock, const int* blockStruct,
                          const double* c, long nnz, const int* k, const int* l, const int* i,
                          const int* j, const double* v, double kappa);

/* out[0..15] = m, nBlock, nSDP, nLP, nDim, packed length, #F1, #F2, #F3,
 * schur_dense (1 if the reference would choose the dense Schur path),
 * bmat_store_first, number of SDP blocks, schur_sparse (1 when the sparse
 * branch is factored by tiles of the ordered matrix: off the dense path and
 * cheaper than the dense factorization), and for it the stored 64 x 64
 * tiles of the factor, the tiles per side and the levels of the tile
 * elimination tree.  Returns the count written. */
int sdpa_gpu_problem_info(sdpa_gpu_ctx* ctx, long* out, int cap);
/* Size of SDP block q (0-based) or -1. */
int sdpa_gpu_block_size(sdpa_gpu_ctx* ctx, int q);

/* Iterate in / out (internal convention, packed block space; y length m).
 * Any pointer may be NULL.  Replaces the Solutions members (sdpa_dataset.h). */
int sdpa_gpu_set_iterate(sdpa_gpu_ctx* ctx, const double* X, const double* y, const double* Z);
int sdpa_gpu_get_iterate(sdpa_gpu_ctx* ctx, double* X, double* y, double* Z);
/* X = Z = lambda I, y = 0, invChol = I/sqrt(lambda), Zinv = I/lambda
 * (Solutions::initialize, sdpa_dataset.cpp:47-74). */
int sdpa_gpu_identity_start(sdpa_gpu_ctx* ctx, double lambda);
/* Remember the current X, Z as the initial point X0, Z0 (init-point runs,
 * sdpa_solve.cpp:61-72). */
int sdpa_gpu_keep_initial_point(sdpa_gpu_ctx* ctx);

/* Solutions::computeInverse (sdpa_dataset.cpp:117-142): Cholesky of every X
 * and Z block (dpotrf semantics), invCholX, invCholZ and Zinv.  Returns 1 if
 * a factorization fails. */
int sdpa_gpu_inv_chol(sdpa_gpu_ctx* ctx);

/* Newton::Make_bMat (sdpa_newton.cpp:2634-2670) into the device bMat (FP64).
 * precision 0: FP64 throughout.  precision 1 (split precision; no
 * counterpart in the reference): the F1/F2 products (G_j = X A_j Z^-1 and
 * A_i . G_j for dense partners) run on FP32 MFMA with FP32 accumulation, over
 * X, Z^-1 and the dense A_i rounded to FP32; their long inner sums are split
 * into slices added in FP64, and the F3 and LP terms stay FP64.  Other values
 * return -30. */
int sdpa_gpu_make_bmat(sdpa_gpu_ctx* ctx, int precision);
/* Lal::choleskyFactorWithAdjust -> rATL_dpotrfL (sdpa_linear.cpp:762-790,
 * sdpa_dpotrf.cpp:218-265): lower Cholesky of bMat with SDPA's pivot rule.
 * Where Chordal::ordering_bMat would leave the dense path (problem_info
 * schur_dense = 0), bMat + 1e-10 I is factored instead, the matrix the
 * reference's sparse branch factors (Chordal::factorizeSchur,
 * sdpa_chordal.cpp:526-534).  When that branch's analysis finds the factor of
 * the ordered matrix cheaper than the dense one (problem_info schur_sparse =
 * 1) the shifted matrix is factored by tiles of the ordered matrix (FP64 at
 * either precision, LAPACK's pivot rule, bMat itself left as assembled;
 * returns a 1-based failing column of the ordered matrix); otherwise the
 * shift is added to the stored bMat once per assembled (or set) matrix and
 * the dense factorization runs.
 * precision 0: FP64, in place.  precision 1 (mixed; no counterpart in the
 * reference): FP32 factor of the column-equilibrated bMat on FP32 MFMA, bMat
 * kept in FP64 for refinement; where that factorization fails or m <= 512 it
 * falls back to FP64 by itself.  Returns 0 or the 1-based failing column. */
int sdpa_gpu_factor_bmat(sdpa_gpu_ctx* ctx, int precision);
/* Lal::solveSystems (sdpa_linear.cpp:843-862): dy = bMat^-1 gVec with the
 * factor left by sdpa_gpu_factor_bmat (the tile factor of the sparse branch:
 * Chordal::solveSchur, sdpa_chordal.cpp:573-579).  After a mixed factorization the
 * solution is refined in FP64 until the residual gVec - bMat dy is at FP64
 * accuracy; returns 2 if the refinement stalls (dy is then not set: factor
 * again with precision 0 and solve again).  The FP64 solve is asynchronous:
 * if its cooperative kernels give up waiting (work-groups not co-resident),
 * the next sdpa_gpu_dots call returns -60 and clears the condition (so does
 * sdpa_gpu_load_problem). */
int sdpa_gpu_solve_bmat(sdpa_gpu_ctx* ctx, int refine);
/* Mixed-precision counters (no counterpart in the reference): out[0]
 * mixed factorizations, [1] fallbacks to FP64 at factorization, [2] refined
 * solves, [3] refinement steps, [4] stalls, [5] 1 if the current factor is
 * the mixed one. */
int sdpa_gpu_mixed_stats(sdpa_gpu_ctx* ctx, long* out6);

/* StepLength::computeStepLength's eigenvalue estimates (sdpa_parts.cpp:363-432,
 * Jal::getMinEigen sdpa_jordan.cpp:58-109): lamX = lambda_min(L_X^-1 Dx L_X^-T),
 * lamZ likewise for Z, each the minimum over blocks (LP included). */
int sdpa_gpu_min_eig(sdpa_gpu_ctx* ctx, double* lamX, double* lamZ);

/* Residuals::compute (sdpa_dataset.cpp:585-608): primal p = b - A(X), dual
 * D = C - sum y_i A_i - Z, and their max norms. */
int sdpa_gpu_residuals(sdpa_gpu_ctx* ctx, double* normPrimalVec, double* normDualMat);
/* Newton::Make_gVec (sdpa_newton.cpp:817-883): r_zinv = target Zinv - X
 * (- Dx Dz Zinv if corrector), minus X D Zinv when `infeasible` (phase noINFO
 * or pFEAS), g = -A(.) + p. */
int sdpa_gpu_make_gvec(sdpa_gpu_ctx* ctx, int corrector, double target, int infeasible);
/* compute_DzMat / compute_DxMat (sdpa_newton.cpp:2718-2752). */
int sdpa_gpu_make_direction(sdpa_gpu_ctx* ctx, int infeasible);
/* Inner products for the control scalars; codes:
 *  0 C.X  1 b.y  2 C.Dx  3 b.dy  4 X.Dz  5 Dx.Z  6 Dx.Dz  7 X.Z
 *  8 trace X  9 trace Z  10 X0.Z  11 X.Z0 */
int sdpa_gpu_dots(sdpa_gpu_ctx* ctx, int nq, const int* codes, double* out);
/* Solutions::update's step (sdpa_dataset.cpp:151-158). */
int sdpa_gpu_update(sdpa_gpu_ctx* ctx, double alphaP, double alphaD);

/* Two calls' work with one host round trip (no counterpart in the reference;
 * for loops that want fewer synchronizations):
 * sdpa_gpu_min_eig_dots = sdpa_gpu_min_eig then sdpa_gpu_dots (nq >= 0 codes)
 * (StepLength::MehrotraCorrector's estimates and the inner products of its
 * rules, sdpa_parts.cpp:493-625, called at sdpa_solve.cpp:155);
 * sdpa_gpu_update_end = sdpa_gpu_inv_chol, sdpa_gpu_dots and
 * sdpa_gpu_residuals after an update (Solutions::update's computeInverse,
 * sdpa_solve.cpp:167, then mu, the residuals and SolveInfo::update,
 * sdpa_solve.cpp:179-190): returns 1 (outputs not set) if an X/Z
 * factorization fails; the inverses are then not meaningful either. */
int sdpa_gpu_min_eig_dots(sdpa_gpu_ctx* ctx, double* lamX, double* lamZ, int nq, const int* codes,
                          double* out);
int sdpa_gpu_update_end(sdpa_gpu_ctx* ctx, int nq, const int* codes, double* out,
                        double* normPrimalVec, double* normDualMat);

/* Buffer access for tests and callers that keep their own copies.
 * which: 0 X, 1 Z, 2 y, 3 Zinv, 4 invCholX, 5 invCholZ, 6 Dx, 7 Dz, 8 dy,
 *        9 gVec, 10 bMat (m*m), 11 primal residual, 12 dual residual.
 * len must equal the buffer's length. */
int sdpa_gpu_get_buffer(sdpa_gpu_ctx* ctx, int which, double* dst, long len);
int sdpa_gpu_set_buffer(sdpa_gpu_ctx* ctx, int which, const double* src, long len);

/* Per-phase device timing (HIP events on the context's stream).  When
 * enabled each timed phase synchronizes.  out[0..9] = ms and call counts for
 * bMat, Cholesky, solve, eigen, inverse. */
int sdpa_gpu_timing(sdpa_gpu_ctx* ctx, int enable, double* out10);
/* Instrumentation with no counterpart in the reference: while timing is
 * enabled, HIP events bracket every launch of the Schur Cholesky's two main
 * kernels -- the wide trailing update on the side stream and the
 * single-launch panel factorization on the context's stream.  out[0..3]:
 * trailing update ms summed over its launches, launches, algorithmic flops
 * (2 K per lower-triangle element updated), 0; out[4..7]: the same for the
 * panel launches (flops: the previous panel's update of the panel, 2 K per
 * element, plus the panel's Cholesky, w^2 h - w^3/3). */
int sdpa_gpu_kernel_timing(sdpa_gpu_ctx* ctx, double* out8);
/* The finer buckets of the same timing, cumulative ms since timing was last
 * enabled (the device-side counterparts of the reference's ComputeTime,
 * sdpa_parts.h:48-86): out[0] bMat, [1] of it the F1/F2 products, [2]
 * Cholesky, [3] solve, [4] eigen (step lengths), [5] inverse Cholesky,
 * [6] gVec, [7] direction (Dz, Dx), [8] residuals, [9] update, [10] inner
 * products.  Returns the count written (at most cap). */
int sdpa_gpu_timing_detail(sdpa_gpu_ctx* ctx, double* out, int cap);
int sdpa_gpu_synchronize(sdpa_gpu_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* SDPA_GPU_H */
