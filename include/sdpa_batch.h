/* sdpa_batch.h -- C ABI of the batched small-SDP solver: many independent
 * single-block instances (n <= 64, m <= 72, sparse constraint matrices)
 * solved by ONE persistent launch, one work-group per instance running
 * SDPA's whole Mehrotra predictor-corrector loop -- Newton direction, step
 * lengths and the control scalars -- on the device.
 *
 * Counterpart in the reference: SDPA::solve (sdpa_solve.cpp:59-233) with
 * its phases (Newton::Mehrotra, StepLength::MehrotraCorrector,
 * Solutions::update, Phase::updateCheck, SolveInfo::update) run once per
 * instance; the formulas are those of the single-problem path (F3 bMat,
 * rATL_dpotrfL pivot rule, Lal::getMinEigen).  Internal convention
 * (reversed primal/dual, C = -F0), default initial point X0 = Z0 =
 * lambda_k I, y0 = 0 with a per-instance lambda_k.
 *
 * Data of instance k (all arrays concatenated over instances):
 *   n[k], m[k]; C: n[k]^2 doubles (column-major, the internal C) at coff[k];
 *   b: m[k] doubles at boff[k]; constraints: entry pointers cptr at
 *   pptr[k] .. pptr[k] + m[k] (m[k] + 1 values, global entry indices), and
 *   entries er, ec (0-based, er <= ec), ev.
 *   At most 1536 stored entries per instance.
 * lambda: the initial-point scale of each instance (> 0).
 * param: maxIteration, epsilonStar, epsilonDash, betaStar, betaBar,
 *   gammaStar, omegaStar, lowerBound, upperBound (9 doubles).
 * threads: 256 or 128 work-items per work-group (one work-group per CU).
 * kernel_ms (may be null): the launch's device time.
 * out: 8 doubles per instance: phase (0 noINFO .. 9 dUNBD, internal
 *   convention), iterations, objP (= C.X), objD (= b.y), normP (max |b -
 *   A(X)|), normD (max |C - sum y A - Z|), mu, failure (0 none, 1 Schur
 *   Cholesky, 2 step length, 3 inverse Cholesky).
 * Returns 0, or < 0: -1 null argument, -2 threads, -3 n or m out of range,
 * -4 offsets out of range, -5 entry pointers, -6 an entry outside the block
 * or below the diagonal, -7 too many entries, -8 lambda <= 0, -100/-101
 * device errors. */
#ifndef SDPA_BATCH_H
#define SDPA_BATCH_H
#ifdef __cplusplus
extern "C" {
#endif

int sdpa_batch_solve(int device, int count, const int* n, const int* m, const long* coff,
                     const double* C, long nC, const long* boff, const double* b, long nb,
                     const long* pptr, const int* cptr, long ncptr, const int* er, const int* ec,
                     const double* ev, long nent, const double* lambda, const double* param,
                     int threads, double* out, double* kernel_ms);

/* Launch configuration of the batched solver.
 *   threads:   work-items per work-group, 128 or 256;
 *   heavy:     F3 pairs (i, j) with more than heavy = |A_i| |A_j| entry
 *              products are evaluated by one wave each (lanes split the
 *              owner's entries), lighter pairs by one thread each (>= 1);
 *   order:     0 instances taken largest first (m^3/3 + n^3), 1 input order;
 *   precision: 0 FP64 Schur Cholesky (rATL_dpotrfL's rule); 1 the Schur
 *              matrix B equilibrated (s_i = B_ii^-1/2) and factored in FP32
 *              (a pivot that is not > 0 falls back to FP64), each solve
 *              refined in FP64 with B until LAPACK dsposv's test holds; a
 *              residual that does not halve switches the instance to the FP64
 *              factorization for the rest of its run;
 *   mu_switch: precision 1: FP64 once mu < mu_switch (0: never).
 * sdpa_batch_solve is sdpa_batch_solve_cfg with {threads, 64, 0, 0, 0}. */
typedef struct {
  int threads;
  int heavy;
  int order;
  int precision;
  double mu_switch;
} sdpa_batch_config;

/* sdpa_batch_solve with a launch configuration.  stats (may be null): 4
 * doubles per instance -- Schur factorizations in FP32, in FP64, refinement
 * steps, refinement stalls.  Returns the codes of sdpa_batch_solve, or -10
 * for a configuration out of range. */
int sdpa_batch_solve_cfg(int device, int count, const int* n, const int* m, const long* coff,
                         const double* C, long nC, const long* boff, const double* b, long nb,
                         const long* pptr, const int* cptr, long ncptr, const int* er,
                         const int* ec, const double* ev, long nent, const double* lambda,
                         const double* param, const sdpa_batch_config* cfg, double* out,
                         double* stats, double* kernel_ms);

/* Autotuner of the launch configuration: solves a sample of the batch
 * (every (count / sample)-th instance, at least 64) under every candidate
 * of threads {256, 128} x heavy {16, 32, 64, 128, 256} x order {0, 1}, at
 * the given precision and mu switch, twice each, and writes the candidate
 * with the smallest kernel time to *best.  table (may be null): 4 doubles
 * per candidate (threads, heavy, order, kernel ms), at most table_cap rows.
 * Returns the number of candidates timed (> 0) or an error code of
 * sdpa_batch_solve_cfg. */
int sdpa_batch_autotune(int device, int count, const int* n, const int* m, const long* coff,
                        const double* C, long nC, const long* boff, const double* b, long nb,
                        const long* pptr, const int* cptr, long ncptr, const int* er,
                        const int* ec, const double* ev, long nent, const double* lambda,
                        const double* param, int precision, double mu_switch, int sample,
                        sdpa_batch_config* best, double* table, int table_cap);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same solve as sdpa_batch_solve, same arguments and
 * results, with no hipMalloc/hipFree and no stream or event creation.  The
 * instance data are staged, and the launch's queue, per-work-group work
 * slabs and results kept, in one fixed device array of the library
 * (sdpa_batch_noalloc_capacity() bytes, 128 MiB, held on each device the
 * library runs on until the process ends).  stream: the HIP stream to run on
 * (NULL: the null stream); the call returns after the results are copied
 * back.  Returns the codes of sdpa_batch_solve, plus -9 when the batch does
 * not fit the array (about 9 MB of fixed slabs per 8 CUs, then roughly
 * 8 n^2 + 16 entries + 100 bytes per instance; split the batch).  Calls that
 * overlap on one device share the array and race: serialize them. */
int sdpa_batch_solve_noalloc(int device, int count, const int* n, const int* m, const long* coff,
                             const double* C, long nC, const long* boff, const double* b, long nb,
                             const long* pptr, const int* cptr, long ncptr, const int* er,
                             const int* ec, const double* ev, long nent, const double* lambda,
                             const double* param, int threads, double* out, void* stream);
long sdpa_batch_noalloc_capacity(void);

#ifdef __cplusplus
}
#endif
#endif /* SDPA_BATCH_H */
