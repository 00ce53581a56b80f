/*
 * sdpa_oracle.c -- CPU FP64 restatement of SDPA 7.4.4's dense-Schur
 * primal-dual interior-point method (Mehrotra predictor-corrector).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (sdpa_amd/) never links it.
 *
 * Every routine cites the reference file:line it restates (paths are
 * relative to the SDPA source tree, LP64/).  The restatement is written
 * from the documented behaviour, not copied; arithmetic follows the same
 * formulas in the same order where the order is observable (pivot rules,
 * loop exits, clipping), and is otherwise plain FP64.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - whole loop: the manual's published example1/example2 traces and
 *     final objectives (manual pp.5, 17-21);
 *   - rATL_dpotrfL, Lal::getMinEigen, Jal::getInvChol/getInvCholAndInv:
 *     against the reference sources compiled under oracle/_ref;
 *   - F1/F2/F3 bMat: closed forms (MaxCut B = X o Z^-1, Lovasz edge
 *     formulas) and the brute-force trace Tr(A_i X A_j Z^-1);
 *   - formula choice: integer-exact restatement, checked on hand cases.
 *
 * Layout conventions (shared with the C-ABI of the HIP library):
 *   "packed block space": SDP blocks column-major, concatenated in internal
 *   block order, followed by the LP coordinates.  bMat is m x m column-major.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))
#define ORC_MIN(a, b) ((a) < (b) ? (a) : (b))

/* ------------------------------------------------------------------ */
/* data structures                                                     */
/* ------------------------------------------------------------------ */

/* One block of one input matrix (SparseMatrix, sdpa_struct.h:82-139). */
typedef struct {
  int n;        /* block dimension */
  int nnz;      /* stored entries (NonZeroCount) */
  int nzeff;    /* NonZeroEffect: diagonal 1, off-diagonal 2 (sdpa_struct.cpp:1338-1342) */
  int *r, *c;   /* 0-based row / column of each stored entry */
  double *v;
  int dense;    /* changeToDense happened (sdpa_struct.cpp:510-555) */
  double *de;   /* n*n column-major when dense */
} osm;

/* SparseLinearSpace (sdpa_struct.h:187-230): present SDP blocks and LP entries. */
typedef struct {
  int nsdp;
  int *sdp_idx; /* internal SDP block number of each present block */
  osm *sdp;
  int nlp;
  int *lp_idx;  /* LP coordinate of each present entry */
  double *lp;
} olin;

typedef struct {
  int m, nBlock;
  int *blockStruct; /* file block sizes, LP negated in the file */
  int *blockType;   /* 1 = SDP, 2 = LP */
  int *blockNumber; /* internal SDP block number or first LP coordinate */
  int nSDP, nLP;
  int *sdpN;        /* SDP block sizes */
  long *sdpOff;     /* packed offsets of SDP blocks */
  long lpOff, total;/* LP offset, total packed length */
  int nDim;         /* sum of SDP sizes + LP count */
  double *b;        /* internal b = file c */
  olin C;           /* internal C = -F0 */
  olin *A;          /* A[0..m-1] = F1..Fm */
  /* inverted index (sdpa_dataset.cpp:268-309) */
  int *sdpNCon; int **sdpCon; int **sdpBlkIdx;
  int *lpNCon; int **lpCon; int **lpBlkIdx;
  /* formula per (constraint, present block) (sdpa_newton.cpp:515-695) */
  int **formula;    /* 1, 2, 3 */
  int calcF1, calcF2;
  int schurDense;   /* Chordal::ordering_bMat decision (sdpa_chordal.cpp:365-410) */
} oprob;

typedef struct {
  int maxIteration;
  double epsilonStar, lambdaStar, omegaStar, lowerBound, upperBound;
  double betaStar, betaBar, gammaStar, epsilonDash;
  double kappa;
} oparam;

enum { noINFO = 0, pFEAS, dFEAS, pdFEAS, pdINF, pFEAS_dINF, pINF_dFEAS, pdOPT, pUNBD, dUNBD };

/* One printed iteration in the internal convention (sdpa_io.cpp:622-658). */
typedef struct {
  double mu, thetaP, thetaD, objP, objD, alphaP, alphaD, beta;
  double bmat_trace; /* diagnostic: sum of bMat diagonal before factorization */
} otrace;

typedef struct {
  int iterations, phase, cholFail, stepFail, invFail;
  double mu, objValPrimal, objValDual, normPrimalVec, normDualMat;
  double time_total, time_bmat, time_chol, time_solve, time_eig, time_inv;
} oresult;

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */

static double now_sec(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* ------------------------------------------------------------------ */
/* problem construction                                                */
/* ------------------------------------------------------------------ */

/* BlockStruct::makeInternalStructure, sdpa_block.cpp:66-108 */
static void make_block_structure(oprob *P) {
  P->blockType = xcalloc(P->nBlock, sizeof(int));
  P->blockNumber = xcalloc(P->nBlock, sizeof(int));
  P->nSDP = 0; P->nLP = 0;
  for (int l = 0; l < P->nBlock; l++) {
    int s = P->blockStruct[l];
    if (s >= 2) { P->blockType[l] = 1; P->blockNumber[l] = P->nSDP++; }
    else if (s < 0) { P->blockType[l] = 2; P->blockNumber[l] = P->nLP; P->nLP += -s; }
    else if (s == 1) { P->blockType[l] = 2; P->blockNumber[l] = P->nLP; P->nLP += 1; }
    else { fprintf(stderr, "oracle: bad block struct\n"); abort(); }
  }
  P->sdpN = xcalloc(P->nSDP, sizeof(int));
  P->sdpOff = xcalloc(P->nSDP + 1, sizeof(long));
  long off = 0; int nd = 0;
  for (int l = 0; l < P->nBlock; l++)
    if (P->blockType[l] == 1) {
      int q = P->blockNumber[l];
      P->sdpN[q] = P->blockStruct[l];
      P->sdpOff[q] = off;
      off += (long)P->blockStruct[l] * P->blockStruct[l];
      nd += P->blockStruct[l];
    }
  P->lpOff = off;
  P->total = off + P->nLP;
  P->nDim = nd + P->nLP;
}

/* SparseMatrix::changeToDense, sdpa_struct.cpp:510-555 (threshold at 518). */
static void change_to_dense(osm *s) {
  if (s->dense) return;
  if ((double)s->nnz < (double)s->n * s->n * 0.20) return;
  int n = s->n;
  s->de = xcalloc((size_t)n * n, sizeof(double));
  for (int e = 0; e < s->nnz; e++) {
    int i = s->r[e], j = s->c[e];
    if (i == j) s->de[i + (size_t)n * j] = s->v[e];
    else s->de[i + (size_t)n * j] = s->de[j + (size_t)n * i] = s->v[e];
  }
  s->dense = 1;
  s->nnz = s->nzeff = n * n;
}

/*
 * Build the problem from .dat-s-style records: k = 0..m (0 is F0),
 * l = 1..nBlock, i, j 1-based.  The record order within one (k, block) is
 * kept.  C is stored negated (sdpa_io.cpp:510-529).  i > j entries are kept
 * as given (the file reader does not swap them; every consumer treats a
 * stored off-diagonal entry symmetrically).
 */
oprob *orc_problem_create(int m, int nBlock, const int *blockStruct, const double *c,
                          long nnz, const int *K, const int *L, const int *I, const int *J,
                          const double *V) {
  oprob *P = xcalloc(1, sizeof(oprob));
  P->m = m; P->nBlock = nBlock;
  P->blockStruct = xcalloc(nBlock, sizeof(int));
  memcpy(P->blockStruct, blockStruct, nBlock * sizeof(int));
  make_block_structure(P);
  P->b = xcalloc(m, sizeof(double));
  memcpy(P->b, c, m * sizeof(double));
  P->A = xcalloc(m, sizeof(olin));

  /* count entries per (k, SDP block) and LP entries per k */
  int nS = P->nSDP;
  int *cnt = xcalloc((size_t)(m + 1) * (nS ? nS : 1), sizeof(int));
  int *lpcnt = xcalloc(m + 1, sizeof(int));
  for (long e = 0; e < nnz; e++) {
    int k = K[e], l = L[e] - 1;
    if (k < 0 || k > m || l < 0 || l >= nBlock) { fprintf(stderr, "oracle: bad record\n"); abort(); }
    if (P->blockType[l] == 1) cnt[(size_t)k * nS + P->blockNumber[l]]++;
    else lpcnt[k]++;
  }
  for (int k = 0; k <= m; k++) {
    olin *T = (k == 0) ? &P->C : &P->A[k - 1];
    int np = 0;
    for (int q = 0; q < nS; q++) if (cnt[(size_t)k * nS + q]) np++;
    T->nsdp = np;
    T->sdp_idx = xcalloc(np, sizeof(int));
    T->sdp = xcalloc(np, sizeof(osm));
    np = 0;
    for (int q = 0; q < nS; q++) {
      int c0 = cnt[(size_t)k * nS + q];
      if (!c0) continue;
      T->sdp_idx[np] = q;
      osm *s = &T->sdp[np];
      s->n = P->sdpN[q];
      s->r = xcalloc(c0, sizeof(int)); s->c = xcalloc(c0, sizeof(int));
      s->v = xcalloc(c0, sizeof(double));
      np++;
    }
    T->lp_idx = xcalloc(lpcnt[k], sizeof(int));
    T->lp = xcalloc(lpcnt[k], sizeof(double));
  }
  for (long e = 0; e < nnz; e++) {
    int k = K[e], l = L[e] - 1;
    olin *T = (k == 0) ? &P->C : &P->A[k - 1];
    double v = (k == 0) ? -V[e] : V[e];
    if (P->blockType[l] == 1) {
      int q = P->blockNumber[l], p;
      for (p = 0; p < T->nsdp; p++) if (T->sdp_idx[p] == q) break;
      osm *s = &T->sdp[p];
      int i = I[e] - 1, j = J[e] - 1;
      if (i < 0 || j < 0 || i >= s->n || j >= s->n) { fprintf(stderr, "oracle: index out of range\n"); abort(); }
      s->r[s->nnz] = i; s->c[s->nnz] = j; s->v[s->nnz] = v; s->nnz++;
      s->nzeff += (i == j) ? 1 : 2; /* sdpa_struct.cpp:1338-1342 */
    } else {
      if (I[e] != J[e]) { fprintf(stderr, "oracle: LP entry must be diagonal\n"); abort(); }
      int idx = P->blockNumber[l] + I[e] - 1, p;
      for (p = 0; p < T->nlp; p++) if (T->lp_idx[p] == idx) break;
      if (p == T->nlp) { T->lp_idx[p] = idx; T->lp[p] = 0.0; T->nlp++; }
      T->lp[p] += v;
    }
  }
  free(cnt); free(lpcnt);
  return P;
}

/* InputData::initialize_index_SDP / _LP, sdpa_dataset.cpp:268-400 */
static void initialize_index(oprob *P) {
  int m = P->m, nS = P->nSDP;
  P->sdpNCon = xcalloc(nS, sizeof(int));
  P->sdpCon = xcalloc(nS, sizeof(int *));
  P->sdpBlkIdx = xcalloc(nS, sizeof(int *));
  for (int k = 0; k < m; k++)
    for (int p = 0; p < P->A[k].nsdp; p++) P->sdpNCon[P->A[k].sdp_idx[p]]++;
  for (int q = 0; q < nS; q++) {
    P->sdpCon[q] = xcalloc(P->sdpNCon[q], sizeof(int));
    P->sdpBlkIdx[q] = xcalloc(P->sdpNCon[q], sizeof(int));
    P->sdpNCon[q] = 0;
  }
  for (int k = 0; k < m; k++)
    for (int p = 0; p < P->A[k].nsdp; p++) {
      int q = P->A[k].sdp_idx[p];
      P->sdpCon[q][P->sdpNCon[q]] = k;
      P->sdpBlkIdx[q][P->sdpNCon[q]] = p;
      P->sdpNCon[q]++;
    }
  int nL = P->nLP;
  P->lpNCon = xcalloc(nL, sizeof(int));
  P->lpCon = xcalloc(nL, sizeof(int *));
  P->lpBlkIdx = xcalloc(nL, sizeof(int *));
  for (int k = 0; k < m; k++)
    for (int p = 0; p < P->A[k].nlp; p++) P->lpNCon[P->A[k].lp_idx[p]]++;
  for (int q = 0; q < nL; q++) {
    P->lpCon[q] = xcalloc(P->lpNCon[q], sizeof(int));
    P->lpBlkIdx[q] = xcalloc(P->lpNCon[q], sizeof(int));
    P->lpNCon[q] = 0;
  }
  for (int k = 0; k < m; k++)
    for (int p = 0; p < P->A[k].nlp; p++) {
      int q = P->A[k].lp_idx[p];
      P->lpCon[q][P->lpNCon[q]] = k;
      P->lpBlkIdx[q][P->lpNCon[q]] = p;
      P->lpNCon[q]++;
    }
}

/* Newton::computeFormula_SDP, sdpa_newton.cpp:515-695 */
static void compute_formula(oprob *P, double kappa) {
  int m = P->m;
  P->formula = xcalloc(m, sizeof(int *));
  int **up = xcalloc(m, sizeof(int *));
  for (int j = 0; j < m; j++) {
    P->formula[j] = xcalloc(P->A[j].nsdp, sizeof(int));
    up[j] = xcalloc(P->A[j].nsdp, sizeof(int));
  }
  /* lines 571-596: count partners with fewer nonzeros, ties with larger index */
  for (int q = 0; q < P->nSDP; q++)
    for (int k1 = 0; k1 < P->sdpNCon[q]; k1++) {
      int j = P->sdpCon[q][k1], jb = P->sdpBlkIdx[q][k1];
      int jnz = P->A[j].sdp[jb].nzeff, u = jnz;
      for (int k2 = 0; k2 < P->sdpNCon[q]; k2++) {
        int i = P->sdpCon[q][k2], ib = P->sdpBlkIdx[q][k2];
        int inz = P->A[i].sdp[ib].nzeff;
        if (inz < jnz) u += inz;
        else if (jnz == inz && i > j) u += inz;
      }
      up[j][jb] = u;
    }
  /* lines 600-676 */
  P->calcF1 = P->calcF2 = 0;
  for (int q = 0; q < P->nSDP; q++)
    for (int k = 0; k < P->sdpNCon[q]; k++) {
      int j = P->sdpCon[q][k], jb = P->sdpBlkIdx[q][k];
      osm *s = &P->A[j].sdp[jb];
      double jnz = s->nzeff, n = s->n, u = up[j][jb];
      double f1 = kappa * n * jnz + n * n * n + kappa * u;
      double f2 = kappa * n * jnz + kappa * (n + 1) * u;
      double f3 = kappa * (2 * kappa * jnz + 1) * u;
      int f;
      if (s->dense) f = (f1 < f2) ? 1 : 2;
      else if (f1 < f2 && f1 < f3) f = (n <= 200 && 2.0 * n >= jnz) ? 3 : 1;
      else if (f2 < f3) f = 2;
      else f = 3;
      P->formula[j][jb] = f;
      if (f == 1) P->calcF1++;
      if (f == 2) P->calcF2++;
    }
  for (int j = 0; j < m; j++) free(up[j]);
  free(up);
}

/* Chordal::ordering_bMat, sdpa_chordal.cpp:365-410 (dense decisions only).
 * Returns 1 when the reference decides DENSE before any MUMPS analysis,
 * 0 when it would go on to the aggregate-sparsity/MUMPS tests. */
static void pattern_mark(unsigned char *pat, int m, const int *g, int ng) {
  for (int a = 0; a < ng; a++)
    for (int b = 0; b <= a; b++) {
      int i = g[a], j = g[b];
      if (i < j) { int t = i; i = j; j = t; }
      pat[(size_t)i * m + j] = 1;
    }
}

/* Chordal::ordering_bMat's dense tests (sdpa_chordal.cpp:365-410): m <= 100;
 * a block or LP coordinate used by more than m sqrt(0.7) constraints; the
 * aggregate pattern (Chordal::makeGraph, 193-300: pairs sharing a block or an
 * LP coordinate, plus the diagonal) holding more than 0.7 m^2 entries.
 * 0 means the reference would go on to the MUMPS analysis. */
static int schur_dense_decision(const oprob *P) {
  if (P->m <= 100) return 1;
  double thr = P->m * sqrt(0.70);
  for (int q = 0; q < P->nSDP; q++) if (P->sdpNCon[q] > thr) return 1;
  for (int q = 0; q < P->nLP; q++) if (P->lpNCon[q] > thr) return 1;
  const int m = P->m;
  unsigned char *pat = (unsigned char *)calloc((size_t)m * m, 1);
  for (int q = 0; q < P->nSDP; q++) pattern_mark(pat, m, P->sdpCon[q], P->sdpNCon[q]);
  for (int q = 0; q < P->nLP; q++) pattern_mark(pat, m, P->lpCon[q], P->lpNCon[q]);
  long lower = 0;
  for (int i = 0; i < m; i++) {
    pat[(size_t)i * m + i] = 1;
    for (int j = 0; j <= i; j++) lower += pat[(size_t)i * m + j];
  }
  free(pat);
  return (double)(2 * lower - m) > 0.70 * (double)m * (double)m;
}

/* SDPA::initializeSolve, sdpa_solve.cpp:10-29 (structural part). */
void orc_problem_prepare(oprob *P, double kappa) {
  for (int p = 0; p < P->C.nsdp; p++) change_to_dense(&P->C.sdp[p]);
  for (int k = 0; k < P->m; k++)
    for (int p = 0; p < P->A[k].nsdp; p++) change_to_dense(&P->A[k].sdp[p]);
  initialize_index(P);
  compute_formula(P, kappa);
  P->schurDense = schur_dense_decision(P);
}

static void free_lin(olin *T) {
  for (int p = 0; p < T->nsdp; p++) {
    free(T->sdp[p].r); free(T->sdp[p].c); free(T->sdp[p].v); free(T->sdp[p].de);
  }
  free(T->sdp); free(T->sdp_idx); free(T->lp); free(T->lp_idx);
}

void orc_problem_free(oprob *P) {
  if (!P) return;
  free_lin(&P->C);
  for (int k = 0; k < P->m; k++) free_lin(&P->A[k]);
  free(P->A);
  if (P->sdpCon) for (int q = 0; q < P->nSDP; q++) { free(P->sdpCon[q]); free(P->sdpBlkIdx[q]); }
  if (P->lpCon) for (int q = 0; q < P->nLP; q++) { free(P->lpCon[q]); free(P->lpBlkIdx[q]); }
  free(P->sdpNCon); free(P->sdpCon); free(P->sdpBlkIdx);
  free(P->lpNCon); free(P->lpCon); free(P->lpBlkIdx);
  if (P->formula) for (int j = 0; j < P->m; j++) free(P->formula[j]);
  free(P->formula);
  free(P->blockStruct); free(P->blockType); free(P->blockNumber);
  free(P->sdpN); free(P->sdpOff); free(P->b);
  free(P);
}

/* queries used by tests */
long orc_packed_size(const oprob *P) { return P->total; }
int orc_n_dim(const oprob *P) { return P->nDim; }
int orc_schur_dense(const oprob *P) { return P->schurDense; }
int orc_formula(const oprob *P, int j, int sdpBlock) {
  for (int p = 0; p < P->A[j].nsdp; p++) if (P->A[j].sdp_idx[p] == sdpBlock) return P->formula[j][p];
  return 0;
}
void orc_formula_counts(const oprob *P, int *out3) {
  out3[0] = out3[1] = out3[2] = 0;
  for (int j = 0; j < P->m; j++)
    for (int p = 0; p < P->A[j].nsdp; p++) out3[P->formula[j][p] - 1]++;
}

/* ------------------------------------------------------------------ */
/* dense kernels on packed block spaces                                 */
/* ------------------------------------------------------------------ */

/* C = A * B for n x n column-major (plain triple loop, j-k-i order). */
static void gemm_nn(int n, const double *A, const double *B, double *C) {
#pragma omp parallel for schedule(static) if (n >= 128)
  for (int j = 0; j < n; j++) {
    double *cj = C + (size_t)n * j;
    for (int i = 0; i < n; i++) cj[i] = 0.0;
    for (int k = 0; k < n; k++) {
      double bkj = B[k + (size_t)n * j];
      const double *ak = A + (size_t)n * k;
      for (int i = 0; i < n; i++) cj[i] += ak[i] * bkj;
    }
  }
}

/* packed triple product R = A*B*C per SDP block; LP entries multiply
 * (Jal::ns_jordan_triple_product, sdpa_jordan.cpp:416-425). */
static void triple_product(const oprob *P, const double *A, const double *B, const double *C,
                           double *R, double *work) {
  for (int q = 0; q < P->nSDP; q++) {
    long o = P->sdpOff[q]; int n = P->sdpN[q];
    gemm_nn(n, A + o, B + o, work + o);
    gemm_nn(n, work + o, C + o, R + o);
  }
  for (int t = 0; t < P->nLP; t++) {
    long o = P->lpOff + t;
    R[o] = A[o] * B[o] * C[o];
  }
}

/* Lal::getSymmetrize, sdpa_linear.cpp:591-618: (a_ij + a_ji) * 0.5 */
static void symmetrize(const oprob *P, double *X) {
  for (int q = 0; q < P->nSDP; q++) {
    long o = P->sdpOff[q]; int n = P->sdpN[q];
    double *a = X + o;
    for (int j = 0; j < n; j++)
      for (int i = j + 1; i < n; i++) {
        double s = (a[i + (size_t)n * j] + a[j + (size_t)n * i]) * 0.5;
        a[i + (size_t)n * j] = a[j + (size_t)n * i] = s;
      }
  }
}

/* dense packed inner product X.Y (DenseLinearSpace '.', sdpa_linear.cpp ddot per block) */
static double dot_packed(const oprob *P, const double *X, const double *Y) {
  double r = 0.0;
  for (int q = 0; q < P->nSDP; q++) {
    long o = P->sdpOff[q], len = (long)P->sdpN[q] * P->sdpN[q];
    double s = 0.0;
    for (long e = 0; e < len; e++) s += X[o + e] * Y[o + e];
    r += s;
  }
  for (int t = 0; t < P->nLP; t++) r += X[P->lpOff + t] * Y[P->lpOff + t];
  return r;
}

/* SparseMatrix . DenseMatrix, sdpa_linear.cpp:318-463 */
static double sparse_dot(const osm *s, const double *D) {
  int n = s->n;
  if (s->dense) {
    double r = 0.0;
    for (long e = 0; e < (long)n * n; e++) r += s->de[e] * D[e];
    return r;
  }
  double r = 0.0;
  for (int e = 0; e < s->nnz; e++) {
    int i = s->r[e], j = s->c[e];
    if (i == j) r += s->v[e] * D[i + (size_t)n * j];
    else r += s->v[e] * (D[i + (size_t)n * j] + D[j + (size_t)n * i]);
  }
  return r;
}

/* SparseLinearSpace . DenseLinearSpace */
static double lin_dot(const oprob *P, const olin *T, const double *D) {
  double r = 0.0;
  for (int p = 0; p < T->nsdp; p++) r += sparse_dot(&T->sdp[p], D + P->sdpOff[T->sdp_idx[p]]);
  for (int p = 0; p < T->nlp; p++) r += T->lp[p] * D[P->lpOff + T->lp_idx[p]];
  return r;
}

/* R += alpha * T (dense R, sparse T), sdpa_linear.cpp:1640-1700 */
static void lin_axpy(const oprob *P, const olin *T, double alpha, double *R) {
  for (int p = 0; p < T->nsdp; p++) {
    const osm *s = &T->sdp[p];
    double *d = R + P->sdpOff[T->sdp_idx[p]];
    int n = s->n;
    if (s->dense) {
      for (long e = 0; e < (long)n * n; e++) d[e] += alpha * s->de[e];
    } else {
      for (int e = 0; e < s->nnz; e++) {
        int i = s->r[e], j = s->c[e];
        double v = s->v[e] * alpha;
        if (i != j) { d[i + (size_t)n * j] += v; d[j + (size_t)n * i] += v; }
        else d[i + (size_t)n * i] += v;
      }
    }
  }
  for (int p = 0; p < T->nlp; p++) R[P->lpOff + T->lp_idx[p]] += alpha * T->lp[p];
}

static double max_abs(const double *x, long n) {
  double r = 0.0;
  for (long e = 0; e < n; e++) { double t = fabs(x[e]); if (t > r) r = t; }
  return r;
}

/* ------------------------------------------------------------------ */
/* Cholesky with SDPA's adjust rule (rATL_dpotrfL, sdpa_dpotrf.cpp)     */
/* ------------------------------------------------------------------ */
#define POTRF_NONZERO 1.0e-14
#define POTRF_ASSIGN 1.0e+100
#define POTRF_LIMIT (-1.0e-6)

/* Unblocked leaf for N <= 4 (potrf4/3/2 and the N==1 case, sdpa_dpotrf.cpp:78-216,
 * 254-262): pivot < -1e-6 fails, pivot < 1e-14 becomes 1e100. */
static int potrf_leaf(int N, double *A, int lda) {
  for (int c = 0; c < N; c++) {
    double d = A[c + (size_t)lda * c];
    for (int k = 0; k < c; k++) d -= A[c + (size_t)lda * k] * A[c + (size_t)lda * k];
    if (d < POTRF_LIMIT) return c + 1;
    if (d < POTRF_NONZERO) d = POTRF_ASSIGN;
    d = sqrt(d);
    A[c + (size_t)lda * c] = d;
    for (int r = c + 1; r < N; r++) {
      double s = A[r + (size_t)lda * c];
      for (int k = 0; k < c; k++) s -= A[r + (size_t)lda * k] * A[c + (size_t)lda * k];
      A[r + (size_t)lda * c] = s / d;
    }
  }
  return 0;
}

/* dtrsm('R','L','T','N', M, N, 1, L, B): B := B * L^-T  (L is N x N lower).
 * Rows are independent: blocks of 64 rows in parallel, columns swept left to
 * right (B[:,c] -= L[c,k] B[:,k] for k < c, then / L[c,c]). */
static void trsm_rltn(int M, int N, const double *L, int lda, double *B, int ldb) {
  const int RB = 64;
  int nrb = (M + RB - 1) / RB;
#pragma omp parallel for schedule(dynamic, 1) if ((long)M * N > 4096)
  for (int rb = 0; rb < nrb; rb++) {
    int r0 = rb * RB, r1 = r0 + RB < M ? r0 + RB : M;
    for (int c = 0; c < N; c++) {
      double *bc = B + (size_t)ldb * c;
      int k = 0;
      for (; k + 4 <= c; k += 4) {
        const double l0 = L[c + (size_t)lda * k], l1 = L[c + (size_t)lda * (k + 1)];
        const double l2 = L[c + (size_t)lda * (k + 2)], l3 = L[c + (size_t)lda * (k + 3)];
        const double *b0 = B + (size_t)ldb * k, *b1 = b0 + ldb, *b2 = b1 + ldb, *b3 = b2 + ldb;
        for (int r = r0; r < r1; r++) bc[r] -= l0 * b0[r] + l1 * b1[r] + l2 * b2[r] + l3 * b3[r];
      }
      for (; k < c; k++) {
        const double lk = L[c + (size_t)lda * k];
        const double *bk = B + (size_t)ldb * k;
        for (int r = r0; r < r1; r++) bc[r] -= lk * bk[r];
      }
      const double d = L[c + (size_t)lda * c];
      for (int r = r0; r < r1; r++) bc[r] /= d;
    }
  }
}

/* dsyrk('L','N', N, K, -1, A, 1, C): lower C -= A A^T, by 64 x 64 tiles of C
 * in parallel, k in blocks of 256, four columns of C per sweep. */
static void syrk_ln(int N, int K, const double *A, int lda, double *C, int ldc) {
  const int TBL = 64, KB = 256;
  int nt = (N + TBL - 1) / TBL;
  long ntiles = (long)nt * (nt + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1) if ((long)N * K > 4096)
  for (long t = 0; t < ntiles; t++) {
    int I = 0;
    long rem = t;
    while (rem >= I + 1) { rem -= I + 1; I++; }
    int J = (int)rem;
    int i0 = I * TBL, i1 = i0 + TBL < N ? i0 + TBL : N;
    int j0 = J * TBL, j1 = j0 + TBL < N ? j0 + TBL : N;
    for (int k0 = 0; k0 < K; k0 += KB) {
      int k1 = k0 + KB < K ? k0 + KB : K;
      int j = j0;
      for (; j + 4 <= j1; j += 4) {
        double *c0 = C + (size_t)ldc * j, *c1 = c0 + ldc, *c2 = c1 + ldc, *c3 = c2 + ldc;
        for (int k = k0; k < k1; k++) {
          const double *ak = A + (size_t)lda * k;
          const double a0 = ak[j], a1 = ak[j + 1], a2 = ak[j + 2], a3 = ak[j + 3];
          int ib = i0 > j ? i0 : j;
          for (int i = ib; i < i1; i++) {
            const double x = ak[i];
            c0[i] -= x * a0;
            if (i >= j + 1) c1[i] -= x * a1;
            if (i >= j + 2) c2[i] -= x * a2;
            if (i >= j + 3) c3[i] -= x * a3;
          }
        }
      }
      for (; j < j1; j++) {
        double *cj = C + (size_t)ldc * j;
        for (int k = k0; k < k1; k++) {
          const double *ak = A + (size_t)lda * k;
          const double ajk = ak[j];
          int ib = i0 > j ? i0 : j;
          for (int i = ib; i < i1; i++) cj[i] -= ak[i] * ajk;
        }
      }
    }
  }
}

/* rATL_dpotrfL, sdpa_dpotrf.cpp:218-265: recursive halving. */
int orc_chol_adjust(int N, double *A, int lda) {
  if (N > 4) {
    int Nl = N >> 1, Nr = N - Nl;
    int ierr = orc_chol_adjust(Nl, A, lda);
    if (ierr) return ierr;
    double *Ar = A + Nl, *An = Ar + (size_t)lda * Nl;
    trsm_rltn(Nr, Nl, A, lda, Ar, lda);
    syrk_ln(Nr, Nl, Ar, lda, An, lda);
    ierr = orc_chol_adjust(Nr, An, lda);
    return ierr ? ierr + Nl : 0;
  }
  return potrf_leaf(N, A, lda);
}

/* Lal::solveSystems, sdpa_linear.cpp:843-862: two dtrsv on the lower factor. */
void orc_chol_solve(int N, const double *L, int lda, double *x) {
  for (int i = 0; i < N; i++) {
    double s = x[i];
    for (int k = 0; k < i; k++) s -= L[i + (size_t)lda * k] * x[k];
    x[i] = s / L[i + (size_t)lda * i];
  }
  for (int i = N - 1; i >= 0; i--) {
    double s = x[i];
    const double *li = L + (size_t)lda * i;
    for (int k = i + 1; k < N; k++) s -= li[k] * x[k];
    x[i] = s / li[i];
  }
}

/* LAPACK dpotrf('L') semantics: fail (info = column) on a non-positive pivot. */
static int potrf_plain(int n, double *A) {
  for (int c = 0; c < n; c++) {
    double d = A[c + (size_t)n * c];
    for (int k = 0; k < c; k++) d -= A[c + (size_t)n * k] * A[c + (size_t)n * k];
    if (!(d > 0.0)) return c + 1;
    d = sqrt(d);
    A[c + (size_t)n * c] = d;
    for (int r = c + 1; r < n; r++) {
      double s = A[r + (size_t)n * c];
      for (int k = 0; k < c; k++) s -= A[r + (size_t)n * k] * A[c + (size_t)n * k];
      A[r + (size_t)n * c] = s / d;
    }
  }
  return 0;
}

/*
 * Jal::getInvChol / getInvCholAndInv, sdpa_jordan.cpp:148-216 and
 * Lal::getCholesky / getInvLowTriangularMatrix, sdpa_linear.cpp:465-515, 566-589:
 * L = chol(A) (upper part zeroed), Linv = L^-1, and optionally Ainv = Linv^T Linv.
 * Returns 0 on success, the failing column otherwise.
 */
int orc_inv_chol(int n, const double *A, double *Linv, double *Ainv, double *Lout) {
  double *L = xcalloc((size_t)n * n, sizeof(double));
  memcpy(L, A, sizeof(double) * n * n);
  int info = potrf_plain(n, L);
  if (info) { free(L); return info; }
  for (int j = 0; j < n; j++) for (int i = 0; i < j; i++) L[i + (size_t)n * j] = 0.0;
  /* Linv = L^-1 by forward substitution on the identity (dtrsm L,L,N,N) */
  memset(Linv, 0, sizeof(double) * n * n);
  for (int j = 0; j < n; j++) {
    double *x = Linv + (size_t)n * j;
    x[j] = 1.0;
    for (int i = j; i < n; i++) {
      double s = x[i];
      for (int k = j; k < i; k++) s -= L[i + (size_t)n * k] * x[k];
      x[i] = s / L[i + (size_t)n * i];
    }
  }
  if (Ainv) {
    /* dtrmm(L, L, T, N): Ainv = Linv^T * Linv */
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        int k0 = i > j ? i : j;
        double s = 0.0;
        for (int k = k0; k < n; k++) s += Linv[k + (size_t)n * i] * Linv[k + (size_t)n * j];
        Ainv[i + (size_t)n * j] = s;
      }
  }
  if (Lout) memcpy(Lout, L, sizeof(double) * n * n);
  free(L);
  return 0;
}

/* Solutions::computeInverse, sdpa_dataset.cpp:117-142 over packed spaces. */
static int compute_inverse(const oprob *P, const double *X, const double *Z,
                           double *invCholX, double *invCholZ, double *Zinv) {
  int ok = 1;
  for (int q = 0; q < P->nSDP; q++) {
    long o = P->sdpOff[q]; int n = P->sdpN[q];
    if (orc_inv_chol(n, X + o, invCholX + o, NULL, NULL)) { ok = 0; break; }
  }
  if (ok)
    for (int t = 0; t < P->nLP; t++) {
      long o = P->lpOff + t;
      if (X[o] < 0.0) { ok = 0; break; }
      invCholX[o] = 1.0 / sqrt(X[o]);
    }
  int okz = 1;
  for (int q = 0; q < P->nSDP; q++) {
    long o = P->sdpOff[q]; int n = P->sdpN[q];
    if (orc_inv_chol(n, Z + o, invCholZ + o, Zinv + o, NULL)) { okz = 0; break; }
  }
  if (okz)
    for (int t = 0; t < P->nLP; t++) {
      long o = P->lpOff + t;
      if (Z[o] < 0.0) { okz = 0; break; }
      invCholZ[o] = 1.0 / sqrt(Z[o]);
      Zinv[o] = 1.0 / Z[o];
    }
  return ok && okz;
}

/* ------------------------------------------------------------------ */
/* eigenvalues: tridiagonal QL (dsteqr 'I' equivalent) and Jacobi       */
/* ------------------------------------------------------------------ */

/*
 * Symmetric tridiagonal eigen-decomposition by the implicit QL method with
 * Wilkinson shifts (Bowdler, Martin, Reinsch & Wilkinson, "The QR and QL
 * algorithms for symmetric matrices", Numer. Math. 11 (1968), EISPACK tql2).
 * d[0..n-1]: diagonal -> eigenvalues ascending; e[0..n-2]: off-diagonal.
 * Z (ldz x n): eigenvectors in columns.  Returns 0, or 1 if an eigenvalue
 * failed to converge in 30*n sweeps (dsteqr's info > 0 case).
 */
int orc_tridiag_ql(int n, double *d, const double *e_in, double *Z, int ldz) {
  double *e = xcalloc(n + 1, sizeof(double));
  for (int i = 0; i < n - 1; i++) e[i] = e_in[i];
  e[n - 1] = 0.0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) Z[i + (size_t)ldz * j] = (i == j) ? 1.0 : 0.0;
  int fail = 0;
  for (int l = 0; l < n && !fail; l++) {
    int iter = 0;
    for (;;) {
      int mm;
      for (mm = l; mm < n - 1; mm++) {
        double dd = fabs(d[mm]) + fabs(d[mm + 1]);
        if (fabs(e[mm]) <= 2.220446049250313e-16 * dd) break;
      }
      if (mm == l) break;
      if (++iter > 30 * n) { fail = 1; break; }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      for (i = mm - 1; i >= l; i--) {
        double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) { d[i + 1] -= p; e[mm] = 0.0; break; }
        s = f / r; c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < n; k++) {
          double zf = Z[k + (size_t)ldz * (i + 1)];
          Z[k + (size_t)ldz * (i + 1)] = s * Z[k + (size_t)ldz * i] + c * zf;
          Z[k + (size_t)ldz * i] = c * Z[k + (size_t)ldz * i] - s * zf;
        }
      }
      if (r == 0.0 && i >= l) continue;
      d[l] -= p; e[l] = g; e[mm] = 0.0;
    }
  }
  /* sort ascending with vectors (selection sort, as dsteqr does) */
  for (int i = 0; i < n - 1; i++) {
    int kmin = i;
    for (int j = i + 1; j < n; j++) if (d[j] < d[kmin]) kmin = j;
    if (kmin != i) {
      double t = d[i]; d[i] = d[kmin]; d[kmin] = t;
      for (int k = 0; k < n; k++) {
        double z = Z[k + (size_t)ldz * i];
        Z[k + (size_t)ldz * i] = Z[k + (size_t)ldz * kmin];
        Z[k + (size_t)ldz * kmin] = z;
      }
    }
  }
  free(e);
  return fail;
}

/* Smallest eigenvalue of a small symmetric matrix (dsyev 'N' equivalent,
 * sdpa_linear.cpp:145-182) by cyclic Jacobi rotations (Rutishauser 1966). */
double orc_sym_min_eig(int n, const double *Ain) {
  double *A = xcalloc((size_t)n * n, sizeof(double));
  /* dsyev('L') reads the lower triangle only */
  for (int j = 0; j < n; j++)
    for (int i = j; i < n; i++) A[i + (size_t)n * j] = A[j + (size_t)n * i] = Ain[i + (size_t)n * j];
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int j = 0; j < n; j++) for (int i = j + 1; i < n; i++) off += A[i + (size_t)n * j] * A[i + (size_t)n * j];
    if (off == 0.0) break;
    double diag = 0.0;
    for (int i = 0; i < n; i++) diag += A[i + (size_t)n * i] * A[i + (size_t)n * i];
    if (off <= 1e-36 * diag) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p + (size_t)n * q];
        if (apq == 0.0) continue;
        double app = A[p + (size_t)n * p], aqq = A[q + (size_t)n * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = A[k + (size_t)n * p], akq = A[k + (size_t)n * q];
          A[k + (size_t)n * p] = c * akp - s * akq;
          A[k + (size_t)n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p + (size_t)n * k], aqk = A[q + (size_t)n * k];
          A[p + (size_t)n * k] = c * apk - s * aqk;
          A[q + (size_t)n * k] = s * apk + c * aqk;
        }
      }
  }
  double mn = A[0];
  for (int i = 1; i < n; i++) if (A[i + (size_t)n * i] < mn) mn = A[i + (size_t)n * i];
  free(A);
  return mn;
}

/*
 * Lal::getMinEigen, sdpa_linear.cpp:29-143: truncated Lanczos estimate of
 * lambda_min(L * D * L^T), L lower triangular, D symmetric, returning
 * min - |error * beta|.  Stop tests at lines 52-56, QL call schedule at 84,
 * result at 128-142.
 */
double orc_lanczos_min(int n, const double *L, const double *D) {
  double *q = xcalloc(n, sizeof(double)), *qold = xcalloc(n, sizeof(double));
  double *r = xcalloc(n, sizeof(double)), *w = xcalloc(n, sizeof(double));
  double *tmp = xcalloc(n, sizeof(double));
  double *diag = xcalloc(n, sizeof(double)), *off = xcalloc(n, sizeof(double));
  double *out = xcalloc(n, sizeof(double)), *bb = xcalloc(n, sizeof(double));
  double *Q = xcalloc((size_t)n * n, sizeof(double));
  double min = 1.0e+51, min_old = 1.0e+52, error = 1.0e+10;
  int k = 0, kk = 0;
  for (int i = 0; i < n; i++) { diag[i] = 1.0e+50; off[i] = 0.0; q[i] = 0.0; r[i] = 1.0; }
  double beta = sqrt((double)n);
  while (k < n && k < sqrt((double)n) + 10 && beta > 1.0e-16 &&
         (fabs(min - min_old) > (1.0e-5) * fabs(min) + (1.0e-8) ||
          fabs(error * beta) > (1.0e-2) * fabs(min) + (1.0e-4))) {
    memcpy(qold, q, sizeof(double) * n);
    double value = 1.0 / beta;
    for (int i = 0; i < n; i++) q[i] = r[i] * value;
    /* w = L^T q (dtrmv Lower, Transpose) */
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = i; j < n; j++) s += L[j + (size_t)n * i] * q[j];
      w[i] = s;
    }
    /* tmp = D w (dgemv) */
    for (int i = 0; i < n; i++) tmp[i] = 0.0;
    for (int j = 0; j < n; j++) {
      double wj = w[j];
      const double *dj = D + (size_t)n * j;
      for (int i = 0; i < n; i++) tmp[i] += dj[i] * wj;
    }
    /* w = L tmp (dtrmv Lower, NoTranspose) */
    for (int i = n - 1; i >= 0; i--) {
      double s = 0.0;
      for (int j = 0; j <= i; j++) s += L[i + (size_t)n * j] * tmp[j];
      w[i] = s;
    }
    double alpha = 0.0;
    for (int i = 0; i < n; i++) alpha += q[i] * w[i];
    diag[k] = alpha;
    for (int i = 0; i < n; i++) r[i] = w[i] - q[i] * alpha;
    for (int i = 0; i < n; i++) r[i] = r[i] - qold[i] * beta;
    if (kk >= sqrt((double)k) || k == n - 1 || k > sqrt((double)n + 9)) {
      kk = 0;
      memcpy(out, diag, sizeof(double) * n);
      memcpy(bb, off, sizeof(double) * n);
      out[n - 1] = diag[k];
      bb[n - 1] = 0.0;
      int info = orc_tridiag_ql(k + 1, out, bb, Q, n);
      if (info > 0) break;
      min_old = min;
      min = out[0];
      error = Q[k];
    }
    double v2 = 0.0;
    for (int i = 0; i < n; i++) v2 += r[i] * r[i];
    beta = sqrt(v2);
    off[k] = beta;
    ++k; ++kk;
  }
  free(q); free(qold); free(r); free(w); free(tmp); free(diag); free(off);
  free(out); free(bb); free(Q);
  return min - fabs(error * beta);
}

/* Jal::getMinEigen(L, X, work), sdpa_jordan.cpp:58-109: minimum over blocks
 * of lambda_min(L X L^T); Lanczos for n > 20, full eigen-solve otherwise. */
double orc_min_eig_packed(const oprob *P, const double *Lp, const double *Dp) {
  double mn = 1.0e50;
  for (int q = 0; q < P->nSDP; q++) {
    long o = P->sdpOff[q]; int n = P->sdpN[q];
    double v;
    if (n > 20) v = orc_lanczos_min(n, Lp + o, Dp + o);
    else {
      double *T1 = xcalloc((size_t)n * n, sizeof(double)), *T2 = xcalloc((size_t)n * n, sizeof(double));
      /* T2 = D * L^T ; T1 = L * T2 */
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          double s = 0.0;
          for (int k = 0; k < n; k++) s += Dp[o + i + (size_t)n * k] * Lp[o + j + (size_t)n * k];
          T2[i + (size_t)n * j] = s;
        }
      gemm_nn(n, Lp + o, T2, T1);
      v = orc_sym_min_eig(n, T1);
      free(T1); free(T2);
    }
    if (v < mn) mn = v;
  }
  for (int t = 0; t < P->nLP; t++) {
    long o = P->lpOff + t;
    double v = Dp[o] * Lp[o] * Lp[o];
    if (v < mn) mn = v;
  }
  return mn;
}

/* ------------------------------------------------------------------ */
/* Schur complement bMat                                               */
/* ------------------------------------------------------------------ */

/* calF3_thread_2, sdpa_newton.cpp:1319-1403: Aj is the owner (outer loop). */
static double calF3(int n, const double *X, const double *Zi, const osm *Ai, const osm *Aj) {
  double ret = 0.0;
  for (int e1 = 0; e1 < Aj->nnz; e1++) {
    int a = Aj->r[e1], b = Aj->c[e1];
    double v1 = Aj->v[e1];
    const double *Xa = X + (size_t)n * a, *Xb = X + (size_t)n * b;
    const double *Za = Zi + (size_t)n * a, *Zb = Zi + (size_t)n * b;
    double sum = 0.0;
    if (a != b) {
      for (int e2 = 0; e2 < Ai->nnz; e2++) {
        int g = Ai->r[e2], d = Ai->c[e2];
        double plu;
        if (g != d) plu = Zb[d] * Xa[g] + Zb[g] * Xa[d] + Za[d] * Xb[g] + Za[g] * Xb[d];
        else plu = Zb[d] * Xa[g] + Za[d] * Xb[g];
        sum += Ai->v[e2] * plu;
      }
    } else {
      for (int e2 = 0; e2 < Ai->nnz; e2++) {
        int g = Ai->r[e2], d = Ai->c[e2];
        double plu;
        if (g != d) plu = Zb[d] * Xa[g] + Zb[g] * Xa[d];
        else plu = Zb[d] * Xa[g];
        sum += Ai->v[e2] * plu;
      }
    }
    ret += v1 * sum;
  }
  return ret;
}

/* X * A_j for sparse or dense A_j (Lal::multiply(Dense,Dense,Sparse), sdpa_linear.cpp:1234-1312) */
static void mult_dense_sparse(int n, const double *X, const osm *A, double *F) {
  if (A->dense) { gemm_nn(n, X, A->de, F); return; }
  memset(F, 0, sizeof(double) * n * n);
  for (int e = 0; e < A->nnz; e++) {
    int i = A->r[e], j = A->c[e];
    double v = A->v[e];
    /* column j of F += v * column i of X ; symmetric partner */
    for (int r = 0; r < n; r++) F[r + (size_t)n * j] += v * X[r + (size_t)n * i];
    if (i != j)
      for (int r = 0; r < n; r++) F[r + (size_t)n * i] += v * X[r + (size_t)n * j];
  }
}

/* Newton::Make_bMat dense path, sdpa_newton.cpp:2634-2670 with
 * compute_bMat_dense_SDP (1827-1922, F1/F2), compute_bMat_dense_SDP3
 * (2019-2125, F3) and compute_bMat_dense_LP (2562-2602). */
void orc_make_bmat(const oprob *P, const double *X, const double *Zinv, double *B) {
  int m = P->m;
  memset(B, 0, sizeof(double) * (size_t)m * m);
  for (int q = 0; q < P->nSDP; q++) {
    long o = P->sdpOff[q]; int n = P->sdpN[q];
    const double *Xq = X + o, *Zq = Zinv + o;
    int nc = P->sdpNCon[q];
    /* F1/F2 pass (only if some F1/F2 exists, line 2650) */
    if (P->calcF1 > 0 || P->calcF2 > 0) {
      double *F = xcalloc((size_t)n * n, sizeof(double)), *G = xcalloc((size_t)n * n, sizeof(double));
      for (int k1 = 0; k1 < nc; k1++) {
        int j = P->sdpCon[q][k1], jb = P->sdpBlkIdx[q][k1];
        int f = P->formula[j][jb];
        if (f == 3) continue;
        const osm *Aj = &P->A[j].sdp[jb];
        int jnz = Aj->nzeff;
        mult_dense_sparse(n, Xq, Aj, F);  /* F = X A_j */
        int hasG = 0;
        if (f == 1) { gemm_nn(n, F, Zq, G); hasG = 1; }  /* G = F Z^-1 */
        for (int k2 = 0; k2 < nc; k2++) {
          int i = P->sdpCon[q][k2], ib = P->sdpBlkIdx[q][k2];
          const osm *Ai = &P->A[i].sdp[ib];
          int inz = Ai->nzeff;
          if (jnz < inz || (inz == jnz && i < j)) continue;
          double value;
          if (f == 1) value = sparse_dot(Ai, G);  /* calF1, 885-889 */
          else {                                  /* calF2, 891-936 */
            if (Ai->dense) {
              if (!hasG) { gemm_nn(n, F, Zq, G); hasG = 1; }
              value = sparse_dot(Ai, G);
            } else {
              value = 0.0;
              for (int e = 0; e < Ai->nnz; e++) {
                int a = Ai->r[e], b = Ai->c[e];
                double s = 0.0;
                for (int t = 0; t < n; t++) s += F[a + (size_t)n * t] * Zq[t + (size_t)n * b];
                value += Ai->v[e] * s;
                if (a != b) {
                  s = 0.0;
                  for (int t = 0; t < n; t++) s += F[b + (size_t)n * t] * Zq[t + (size_t)n * a];
                  value += Ai->v[e] * s;
                }
              }
            }
          }
          if (i != j) { B[i + (size_t)m * j] += value; B[j + (size_t)m * i] += value; }
          else B[i + (size_t)m * i] += value;
        }
      }
      free(F); free(G);
    }
    /* F3 pass */
#pragma omp parallel for schedule(dynamic, 16)
    for (int k1 = 0; k1 < nc; k1++) {
      int j = P->sdpCon[q][k1], jb = P->sdpBlkIdx[q][k1];
      if (P->formula[j][jb] != 3) continue;
      const osm *Aj = &P->A[j].sdp[jb];
      int jnz = Aj->nzeff;
      for (int k2 = 0; k2 < nc; k2++) {
        int i = P->sdpCon[q][k2], ib = P->sdpBlkIdx[q][k2];
        const osm *Ai = &P->A[i].sdp[ib];
        int inz = Ai->nzeff;
        if (jnz < inz || (inz == jnz && i < j)) continue;
        double value = calF3(n, Xq, Zq, Ai, Aj);
        if (i != j) { B[i + (size_t)m * j] += value; B[j + (size_t)m * i] += value; }
        else B[i + (size_t)m * i] += value;
      }
    }
  }
  /* LP pass */
  for (int t = 0; t < P->nLP; t++) {
    double x = X[P->lpOff + t], zi = Zinv[P->lpOff + t];
    for (int k1 = 0; k1 < P->lpNCon[t]; k1++) {
      int j = P->lpCon[t][k1];
      double aj = P->A[j].lp[P->lpBlkIdx[t][k1]];
      for (int k2 = k1; k2 < P->lpNCon[t]; k2++) {
        int i = P->lpCon[t][k2];
        double ai = P->A[i].lp[P->lpBlkIdx[t][k2]];
        double value = x * zi * ai * aj;
        if (i != j) { B[i + (size_t)m * j] += value; B[j + (size_t)m * i] += value; }
        else B[i + (size_t)m * i] += value;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* the interior-point loop                                             */
/* ------------------------------------------------------------------ */

typedef struct {
  const oprob *P;
  oparam prm;
  long T; int m, nDim;
  double *X, *Z, *y, *invCholX, *invCholZ, *Zinv;
  double *Dx, *Dz, *dy, *rzinv, *xrdzinv, *g, *work1, *work2;
  double *primalVec, *dualMat;       /* currentRes */
  double initNormP, initNormD, normP, normD;
  double *B;
  /* control scalars */
  double muInit, muCur, thetaP, thetaD, betaVal, alphaP, alphaD;
  double rho, etaP, etaD, objP, objD;
  int phase, affine;
  int isInitPoint;
  double *X0, *Z0;
} ostate;

/* Chordal::factorizeSchur (sdpa_chordal.cpp:526-534): when ordering_bMat
 * leaves the dense path, the reference factors bMat + PLUS_ADJUST_DIAGONAL I
 * (1e-10, sdpa_chordal.cpp:25) by its sparse solver; the same shifted matrix
 * is factored here (densely). */
static void schur_shift(const oprob *P, double *B) {
  if (P->schurDense) return;
  for (int i = 0; i < P->m; i++) B[i + (size_t)P->m * i] += 1.0e-10;
}

/* Residuals::compute, sdpa_dataset.cpp:585-608 */
static void residuals(ostate *S) {
  const oprob *P = S->P;
  for (int i = 0; i < S->m; i++) S->primalVec[i] = -lin_dot(P, &P->A[i], S->X) + P->b[i];
  memset(S->dualMat, 0, sizeof(double) * S->T);
  for (int i = 0; i < S->m; i++) lin_axpy(P, &P->A[i], S->y[i], S->dualMat);
  for (long e = 0; e < S->T; e++) S->dualMat[e] = -S->dualMat[e];
  lin_axpy(P, &P->C, 1.0, S->dualMat);
  for (long e = 0; e < S->T; e++) S->dualMat[e] -= S->Z[e];
  S->normP = max_abs(S->primalVec, S->m);
  S->normD = max_abs(S->dualMat, S->T);
}

/* Phase::updateCheck, sdpa_parts.cpp:1150-1242 (REVERSE_PRIMAL_DUAL=1) */
static int phase_update(ostate *S) {
  const double NONZERO = 1.0e-6;
  double acc = S->prm.epsilonDash;
  int v = noINFO;
  if (S->normP <= acc) v = (S->normD <= acc) ? pdFEAS : pFEAS;
  if (v == noINFO && S->normD <= acc) v = dFEAS;
  S->phase = v;
  if (v == pdFEAS) {
    double mean = (fabs(S->objP) + fabs(S->objD)) / 2.0;
    double gap = fabs(S->objP - S->objD);
    double dom = mean < 1.0 ? 1.0 : mean;
    if (gap / dom <= S->prm.epsilonStar) { S->phase = pdOPT; return 0; }
  }
  if (v == noINFO && S->rho > 1.0 + NONZERO) { S->phase = pdINF; return 0; }
  if (v == pFEAS) {
    if (S->objP <= -S->prm.upperBound) { S->phase = pUNBD; return 0; }
    if (S->rho > 1.0 + NONZERO) { S->phase = pFEAS_dINF; return 0; }
  }
  if (v == dFEAS) {
    if (S->objD >= -S->prm.lowerBound) { S->phase = dUNBD; return 0; }
    if (S->rho > 1.0 + NONZERO) { S->phase = pINF_dFEAS; return 0; }
  }
  return 1;
}

/* Newton::Make_gVec / compute_rMat, sdpa_newton.cpp:793-883 */
static void make_gvec(ostate *S, int corrector) {
  const oprob *P = S->P;
  double target = S->betaVal * S->muCur;
  for (long e = 0; e < S->T; e++) S->rzinv[e] = S->Zinv[e] * target;
  for (long e = 0; e < S->T; e++) S->rzinv[e] = S->rzinv[e] + S->X[e] * (-1.0);
  if (corrector) {
    triple_product(P, S->Dx, S->Dz, S->Zinv, S->work1, S->work2);
    for (long e = 0; e < S->T; e++) S->rzinv[e] = S->rzinv[e] + S->work1[e] * (-1.0);
  }
  if (S->phase == pFEAS || S->phase == noINFO) {
    if (!corrector) triple_product(P, S->X, S->dualMat, S->Zinv, S->xrdzinv, S->work2);
    for (long e = 0; e < S->T; e++) S->work1[e] = S->rzinv[e] + S->xrdzinv[e] * (-1.0);
  } else {
    memcpy(S->work1, S->rzinv, sizeof(double) * S->T);
  }
  for (int i = 0; i < S->m; i++) S->g[i] = -lin_dot(P, &P->A[i], S->work1) + S->primalVec[i];
}

/* compute_DzMat / compute_DxMat, sdpa_newton.cpp:2718-2752 */
static void make_dz_dx(ostate *S) {
  const oprob *P = S->P;
  memset(S->Dz, 0, sizeof(double) * S->T);
  for (int i = 0; i < S->m; i++) lin_axpy(P, &P->A[i], S->dy[i], S->Dz);
  for (long e = 0; e < S->T; e++) S->Dz[e] = -S->Dz[e];
  if (S->phase == pFEAS || S->phase == noINFO)
    for (long e = 0; e < S->T; e++) S->Dz[e] += S->dualMat[e];
  triple_product(P, S->X, S->Dz, S->Zinv, S->work1, S->work2);
  for (long e = 0; e < S->T; e++) S->Dx[e] = S->rzinv[e] + S->work1[e] * (-1.0);
  symmetrize(P, S->Dx);
}

/* StepLength::MehrotraPredictor, sdpa_parts.cpp:435-491 */
static void alpha_predictor(ostate *S) {
  const oprob *P = S->P;
  S->alphaP = S->alphaD = 0.9;
  if (S->phase == noINFO || S->phase == dFEAS) {
    if (S->alphaP > 1.0) S->alphaP = 1.0;
  } else {
    double inc = lin_dot(P, &P->C, S->Dx);
    if (inc > 0.0) {
      if (S->alphaP > S->alphaD) S->alphaP = S->alphaD;
      if (S->alphaP > 1.0) S->alphaP = 1.0;
    }
  }
  if (S->phase == noINFO || S->phase == pFEAS) {
    if (S->alphaD > 1.0) S->alphaD = 1.0;
  } else {
    double inc = 0.0;
    for (int i = 0; i < S->m; i++) inc += P->b[i] * S->dy[i];
    if (inc < 0.0) {
      if (S->alphaD > S->alphaP) S->alphaD = S->alphaP;
      if (S->alphaD > 1.0) S->alphaD = 1.0;
    }
  }
}

/* DirectionParameter::MehrotraCorrector, sdpa_parts.cpp:668-706 */
static void beta_corrector(ostate *S) {
  const oprob *P = S->P;
  double xv = dot_packed(P, S->X, S->Dz), uz = dot_packed(P, S->Dx, S->Z), uv = dot_packed(P, S->Dx, S->Dz);
  double muT = S->muCur + (S->alphaP * uz + S->alphaD * xv + S->alphaP * S->alphaD * uv) / S->nDim;
  double v = muT / S->muCur;
  if (v < 1.0) v = v * v;
  if (S->phase == pdFEAS) {
    if (v < S->prm.betaStar) v = S->prm.betaStar;
    if (v > 1.0) v = 1.0;
  } else {
    if (v < S->prm.betaBar) v = S->prm.betaBar;
  }
  S->betaVal = v;
}

/* StepLength::computeStepLength + MehrotraCorrector, sdpa_parts.cpp:363-625 */
static void alpha_corrector(ostate *S, double *teig) {
  const oprob *P = S->P;
  double t0 = now_sec();
  const double alphaBD = 100.0;
  double mx = orc_min_eig_packed(P, S->invCholX, S->Dx);
  S->alphaP = (-mx > 1.0 / alphaBD) ? -1.0 / mx : alphaBD;
  double mz = orc_min_eig_packed(P, S->invCholZ, S->Dz);
  S->alphaD = (-mz > 1.0 / alphaBD) ? -1.0 / mz : alphaBD;
  *teig += now_sec() - t0;
  S->alphaP *= S->prm.gammaStar;
  S->alphaD *= S->prm.gammaStar;
  if (S->phase == noINFO || S->phase == dFEAS) {
    if (S->alphaP > 1.0) S->alphaP = 1.0;
  } else {
    double inc = lin_dot(P, &P->C, S->Dx);
    if (inc > 0.0) {
      if (S->alphaP > S->alphaD) S->alphaP = S->alphaD;
      if (S->alphaP > 1.0) S->alphaP = 1.0;
    }
  }
  if (S->phase == noINFO || S->phase == pFEAS) {
    if (S->alphaD > 1.0) S->alphaD = 1.0;
  } else {
    double inc = 0.0;
    for (int i = 0; i < S->m; i++) inc += P->b[i] * S->dy[i];
    if (inc < 0.0) {
      if (S->alphaD > S->alphaP) S->alphaD = S->alphaP;
      if (S->alphaD > 1.0) S->alphaD = 1.0;
    }
  }
  /* lines 565-598: attain feasibility before mu reduction */
  if (!S->affine && (S->phase == noINFO || S->phase == pFEAS || S->phase == dFEAS)) {
    double xv = dot_packed(P, S->X, S->Dz), uz = dot_packed(P, S->Dx, S->Z), uv = dot_packed(P, S->Dx, S->Dz);
    const double xi = 3.0;
    double thetaMax = ORC_MAX((1.0 - S->alphaP) * S->thetaP, (1.0 - S->alphaD) * S->thetaD);
    double muNew = S->muCur + (S->alphaP * uz + S->alphaD * xv + S->alphaP * S->alphaD * uv) / S->nDim;
    while (thetaMax * thetaMax * S->muInit > xi * muNew) {
      double aMax = 0.95 * ORC_MAX(S->alphaP, S->alphaD);
      S->alphaP = ORC_MIN(S->alphaP, aMax);
      S->alphaD = ORC_MIN(S->alphaD, aMax);
      thetaMax = ORC_MAX((1.0 - S->alphaP) * S->thetaP, (1.0 - S->alphaD) * S->thetaD);
      muNew = S->muCur + (S->alphaP * uz + S->alphaD * xv + S->alphaP * S->alphaD * uv) / S->nDim;
      if (S->alphaP < 1.0e-6 && S->alphaD < 1.0e-6) break;
    }
  }
  /* lines 602-624: keep primal objective >= dual objective */
  if (S->phase == pdFEAS) {
    double objD = 0.0, incD = 0.0;
    for (int i = 0; i < S->m; i++) { objD += P->b[i] * S->y[i]; incD += P->b[i] * S->dy[i]; }
    double objP = lin_dot(P, &P->C, S->X);
    incD *= S->alphaD;
    double incP = lin_dot(P, &P->C, S->Dx) * S->alphaP;
    double maxRatio = (objD - objP) / (incP - incD);
    if (maxRatio > 0.0 && maxRatio < 1.0) { S->alphaP *= maxRatio; S->alphaD *= maxRatio; }
  }
}

/* SolveInfo::update (both variants), sdpa_parts.cpp:914-1060 */
static void solveinfo_update(ostate *S) {
  const oprob *P = S->P;
  S->objP = lin_dot(P, &P->C, S->X);
  double od = 0.0;
  for (int i = 0; i < S->m; i++) od += P->b[i] * S->y[i];
  S->objD = od;
  double primal = S->thetaP, dual = S->thetaD, omega = S->prm.omegaStar;
  S->rho = 0.0;
  double x0z0 = S->nDim * S->muInit, xz = S->nDim * S->muCur;
  double x0z = 0.0, xz0 = 0.0;
  if (S->isInitPoint) {
    x0z = dot_packed(P, S->X0, S->Z);
    xz0 = dot_packed(P, S->X, S->Z0);
  } else {
    double lam = S->prm.lambdaStar;
    for (int q = 0; q < P->nSDP; q++) {
      long o = P->sdpOff[q]; int n = P->sdpN[q];
      for (int k = 0; k < n; k++) {
        x0z += lam * S->Z[o + (long)k * n + k];
        xz0 += lam * S->X[o + (long)k * n + k];
      }
    }
    for (int t = 0; t < P->nLP; t++) { x0z += lam * S->Z[P->lpOff + t]; xz0 += lam * S->X[P->lpOff + t]; }
  }
  double acc = S->prm.epsilonDash;
  if (S->normP <= acc && xz0 < S->etaP) S->etaP = xz0;
  if (S->normD <= acc && x0z < S->etaD) S->etaD = x0z;
  if (S->normP > acc && S->normD <= acc)
    S->rho = primal * x0z / ((primal + (1.0 - primal) * omega) * S->etaD + xz);
  if (S->normP <= acc && S->normD > acc)
    S->rho = dual * xz0 / ((dual + (1.0 - dual) * omega) * S->etaP + xz);
  if (S->normP > acc && S->normD > acc)
    S->rho = (dual * xz0 + primal * x0z) /
             ((primal * dual + omega * (primal * (1.0 - dual) + (1.0 - primal) * dual)) * x0z0 + xz);
}

void orc_default_param(oparam *p) {
  /* Parameter::setDefaultParameter, sdpa_parts.cpp:208-219 */
  p->maxIteration = 100; p->epsilonStar = 1.0e-7; p->lambdaStar = 1.0e+2;
  p->omegaStar = 2.0; p->lowerBound = -1.0e+5; p->upperBound = 1.0e+5;
  p->betaStar = 0.1; p->betaBar = 0.2; p->gammaStar = 0.9; p->epsilonDash = 1.0e-7;
  p->kappa = 12.0; /* sdpa_call.cpp:41 */
}

/*
 * SDPA::initializeSolve + SDPA::solve, sdpa_solve.cpp:10-233.
 * X0/Y0 (internal X, Z packed) and y0 may be NULL (default lambda* I start).
 * trace must hold maxIteration+1 entries (the last one is the final line
 * printed by printLastInfo).  outX/outZ/outY may be NULL.
 * The problem must have been prepared (orc_problem_prepare).
 */
int orc_solve(const oprob *P, const oparam *prm, const double *X0, const double *Z0,
              const double *y0, otrace *trace, oresult *res, double *outX, double *outZ,
              double *outY, int maxIterOverride) {
  ostate s; memset(&s, 0, sizeof(s));
  ostate *S = &s;
  double tstart = now_sec();
  S->P = P; S->prm = *prm; S->T = P->total; S->m = P->m; S->nDim = P->nDim;
  if (maxIterOverride > 0) S->prm.maxIteration = maxIterOverride;
  long T = S->T; int m = S->m;
  double **bufs[] = {&S->X, &S->Z, &S->invCholX, &S->invCholZ, &S->Zinv, &S->Dx, &S->Dz,
                     &S->rzinv, &S->xrdzinv, &S->work1, &S->work2, &S->dualMat};
  for (unsigned b = 0; b < sizeof(bufs) / sizeof(bufs[0]); b++) *bufs[b] = xcalloc(T, sizeof(double));
  S->y = xcalloc(m, sizeof(double)); S->dy = xcalloc(m, sizeof(double));
  S->g = xcalloc(m, sizeof(double)); S->primalVec = xcalloc(m, sizeof(double));
  S->B = xcalloc((size_t)m * m, sizeof(double));
  double lam = S->prm.lambdaStar;
  memset(res, 0, sizeof(*res));
  S->isInitPoint = (X0 != NULL);
  if (!S->isInitPoint) {
    /* Solutions::initialize, sdpa_dataset.cpp:47-74 */
    for (int q = 0; q < P->nSDP; q++) {
      long o = P->sdpOff[q]; int n = P->sdpN[q];
      for (int k = 0; k < n; k++) {
        long d = o + (long)k * n + k;
        S->X[d] = lam; S->Z[d] = lam;
        S->invCholX[d] = 1.0 / sqrt(lam); S->invCholZ[d] = 1.0 / sqrt(lam);
        S->Zinv[d] = 1.0 / lam;
      }
    }
    for (int t = 0; t < P->nLP; t++) {
      long d = P->lpOff + t;
      S->X[d] = lam; S->Z[d] = lam;
      S->invCholX[d] = 1.0 / sqrt(lam); S->invCholZ[d] = 1.0 / sqrt(lam);
      S->Zinv[d] = 1.0 / lam;
    }
    S->muInit = S->muCur = lam * lam;
  } else {
    memcpy(S->X, X0, sizeof(double) * T);
    memcpy(S->Z, Z0, sizeof(double) * T);
    if (y0) memcpy(S->y, y0, sizeof(double) * m);
    S->muInit = S->muCur = dot_packed(P, S->X, S->Z) / S->nDim;
    if (!compute_inverse(P, S->X, S->Z, S->invCholX, S->invCholZ, S->Zinv)) res->invFail = 1;
    S->X0 = xcalloc(T, sizeof(double)); S->Z0 = xcalloc(T, sizeof(double));
    memcpy(S->X0, S->X, sizeof(double) * T); memcpy(S->Z0, S->Z, sizeof(double) * T);
  }
  residuals(S);
  S->initNormP = S->normP; S->initNormD = S->normD;
  S->betaVal = S->prm.betaStar;
  S->thetaP = (S->initNormP < S->prm.epsilonDash) ? 0.0 : 1.0;
  S->thetaD = (S->initNormD < S->prm.epsilonDash) ? 0.0 : 1.0;
  S->rho = 1.0;
  S->etaP = S->etaD = S->prm.omegaStar * S->nDim * S->muInit;
  S->objP = lin_dot(P, &P->C, S->X);
  { double od = 0.0; for (int i = 0; i < m; i++) od += P->b[i] * S->y[i]; S->objD = od; }
  phase_update(S);

  int it = 0;
  while (!res->invFail && phase_update(S) && it < S->prm.maxIteration) {
    /* Switch::MehrotraPredictor, sdpa_parts.cpp:733-743 */
    S->affine = !(S->phase == noINFO || S->phase == pFEAS || S->phase == dFEAS);
    /* DirectionParameter::MehrotraPredictor, 653-666 */
    if (S->phase == pdFEAS) S->betaVal = 0.0;
    else { S->betaVal = S->prm.betaBar; if (S->affine) S->betaVal = 2.0; }

    /* Newton::Mehrotra(PREDICTOR), sdpa_newton.cpp:2755-2809 */
    make_gvec(S, 0);
    double t0 = now_sec();
    orc_make_bmat(P, S->X, S->Zinv, S->B);
    double t1 = now_sec();
    res->time_bmat += t1 - t0;
    double tr = 0.0;
    for (int i = 0; i < m; i++) tr += S->B[i + (size_t)m * i];
    schur_shift(P, S->B);
    int info = orc_chol_adjust(m, S->B, m);
    double t2 = now_sec();
    res->time_chol += t2 - t1;
    if (info) { res->cholFail = info; break; }
    memcpy(S->dy, S->g, sizeof(double) * m);
    orc_chol_solve(m, S->B, m, S->dy);
    res->time_solve += now_sec() - t2;
    make_dz_dx(S);
    alpha_predictor(S);
    beta_corrector(S);
    /* Newton::Mehrotra(CORRECTOR) */
    make_gvec(S, 1);
    t2 = now_sec();
    memcpy(S->dy, S->g, sizeof(double) * m);
    orc_chol_solve(m, S->B, m, S->dy);
    res->time_solve += now_sec() - t2;
    make_dz_dx(S);
    alpha_corrector(S, &res->time_eig);

    if (trace) {
      otrace *tr0 = &trace[it];
      tr0->mu = S->muCur; tr0->thetaP = S->thetaP; tr0->thetaD = S->thetaD;
      tr0->objP = S->objP; tr0->objD = S->objD; tr0->alphaP = S->alphaP;
      tr0->alphaD = S->alphaD; tr0->beta = S->betaVal; tr0->bmat_trace = tr;
    }
    /* Solutions::update, sdpa_dataset.cpp:145-171 */
    for (long e = 0; e < T; e++) S->X[e] = S->X[e] + S->Dx[e] * S->alphaP;
    for (int i = 0; i < m; i++) S->y[i] = S->y[i] + S->dy[i] * S->alphaD;
    for (long e = 0; e < T; e++) S->Z[e] = S->Z[e] + S->Dz[e] * S->alphaD;
    if (S->alphaP < 1.0e-4 && S->alphaD < 1.0e-4) { res->stepFail = 1; break; }
    t2 = now_sec();
    int okinv = compute_inverse(P, S->X, S->Z, S->invCholX, S->invCholZ, S->Zinv);
    res->time_inv += now_sec() - t2;
    if (!okinv) { res->invFail = 1; break; }
    /* RatioInitResCurrentRes::update, 838-846 */
    if (!S->affine) {
      S->thetaP = fabs((1.0 - S->alphaP) * S->thetaP);
      S->thetaD = fabs((1.0 - S->alphaD) * S->thetaD);
    }
    S->muCur = dot_packed(P, S->X, S->Z) / S->nDim;
    residuals(S);
    /* update_exact, 848-866 */
    S->thetaP = (S->initNormP > S->prm.epsilonDash * 1.0e-2) ? S->normP / S->initNormP : 0.0;
    S->thetaD = (S->initNormD > S->prm.epsilonDash * 1.0e-2) ? S->normD / S->initNormD : 0.0;
    solveinfo_update(S);
    it++;
  }
  residuals(S);
  if (trace) {
    otrace *tr0 = &trace[it];
    tr0->mu = S->muCur; tr0->thetaP = S->thetaP; tr0->thetaD = S->thetaD;
    tr0->objP = S->objP; tr0->objD = S->objD; tr0->alphaP = S->alphaP;
    tr0->alphaD = S->alphaD; tr0->beta = S->betaVal; tr0->bmat_trace = 0.0;
  }
  res->iterations = it;
  res->phase = S->phase;
  res->mu = S->muCur;
  res->objValPrimal = S->objP; res->objValDual = S->objD;
  res->normPrimalVec = S->normP; res->normDualMat = S->normD;
  res->time_total = now_sec() - tstart;
  if (outX) memcpy(outX, S->X, sizeof(double) * T);
  if (outZ) memcpy(outZ, S->Z, sizeof(double) * T);
  if (outY) memcpy(outY, S->y, sizeof(double) * m); /* internal sign; caller negates */
  for (unsigned b = 0; b < sizeof(bufs) / sizeof(bufs[0]); b++) free(*bufs[b]);
  free(S->y); free(S->dy); free(S->g); free(S->primalVec); free(S->B);
  free(S->X0); free(S->Z0);
  return it;
}

/*
 * One Mehrotra Newton step from a given iterate, without the step lengths:
 * Newton::Mehrotra(PREDICTOR) then Newton::Mehrotra(CORRECTOR) with the
 * predictor's factor reused (sdpa_newton.cpp:2755-2809; Make_gVec 793-883,
 * compute_DyVec 2672-2716, compute_DzMat/DxMat 2718-2752).  The iterate
 * (X, y, Z internal, packed) is taken as is; computeInverse runs first
 * (sdpa_dataset.cpp:117-142).  infeasible != 0 takes the branch of phases
 * noINFO/pFEAS (the X D Z^-1 terms and D added to Dz).  The predictor's
 * target is betaPred * mu, the corrector's betaCorr * mu.
 * out (length 2 * (2 m + 2 T)): predictor g, dy, Dx, Dz, then the
 * corrector's.  Returns 0, the failing Schur pivot (> 0), or -1 if X or Z
 * is not positive definite.
 */
int orc_direction(const oprob *P, const double *X, const double *y, const double *Z,
                  int infeasible, double mu, double betaPred, double betaCorr, double *out) {
  ostate s; memset(&s, 0, sizeof(s));
  ostate *S = &s;
  S->P = P; S->T = P->total; S->m = P->m; S->nDim = P->nDim;
  long T = S->T; int m = S->m;
  double **bufs[] = {&S->X, &S->Z, &S->invCholX, &S->invCholZ, &S->Zinv, &S->Dx, &S->Dz,
                     &S->rzinv, &S->xrdzinv, &S->work1, &S->work2, &S->dualMat};
  for (unsigned b = 0; b < sizeof(bufs) / sizeof(bufs[0]); b++) *bufs[b] = xcalloc(T, sizeof(double));
  S->y = xcalloc(m, sizeof(double)); S->dy = xcalloc(m, sizeof(double));
  S->g = xcalloc(m, sizeof(double)); S->primalVec = xcalloc(m, sizeof(double));
  S->B = xcalloc((size_t)m * m, sizeof(double));
  memcpy(S->X, X, sizeof(double) * T);
  memcpy(S->Z, Z, sizeof(double) * T);
  memcpy(S->y, y, sizeof(double) * m);
  int rc = 0;
  if (!compute_inverse(P, S->X, S->Z, S->invCholX, S->invCholZ, S->Zinv)) { rc = -1; goto done; }
  residuals(S);
  S->phase = infeasible ? noINFO : pdFEAS;
  S->muCur = mu;
  S->betaVal = betaPred;
  make_gvec(S, 0);
  orc_make_bmat(P, S->X, S->Zinv, S->B);
  schur_shift(P, S->B);
  rc = orc_chol_adjust(m, S->B, m);
  if (rc) goto done;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) { S->betaVal = betaCorr; make_gvec(S, 1); }
    memcpy(S->dy, S->g, sizeof(double) * m);
    orc_chol_solve(m, S->B, m, S->dy);
    make_dz_dx(S);
    double *o = out + (size_t)pass * (2 * m + 2 * T);
    memcpy(o, S->g, sizeof(double) * m);
    memcpy(o + m, S->dy, sizeof(double) * m);
    memcpy(o + 2 * m, S->Dx, sizeof(double) * T);
    memcpy(o + 2 * m + T, S->Dz, sizeof(double) * T);
  }
done:
  for (unsigned b = 0; b < sizeof(bufs) / sizeof(bufs[0]); b++) free(*bufs[b]);
  free(S->y); free(S->dy); free(S->g); free(S->primalVec); free(S->B);
  return rc;
}

/* Size of the structs for the ctypes mirror in tests. */
int orc_sizeof_param(void) { return (int)sizeof(oparam); }
int orc_sizeof_trace(void) { return (int)sizeof(otrace); }
int orc_sizeof_result(void) { return (int)sizeof(oresult); }
int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
