// batch.hip -- many small SDPs solved by one persistent launch.  One
// work-group per instance runs SDPA's whole Mehrotra predictor-corrector
// loop (SDPA::solve, sdpa_solve.cpp:59-233): the Newton direction (Make_gVec
// sdpa_newton.cpp:817-883, F3 bMat 1319-1403, rATL_dpotrfL sdpa_dpotrf.cpp:
// 218-265, compute_DzMat/DxMat 2718-2752), the step lengths (Lal::
// getMinEigen, sdpa_linear.cpp:29-182) and the control scalars
// (sdpa_parts.cpp:493-1242) -- the same formulas as the single-problem path,
// with every scalar decision taken on the device.  Instances: one SDP block
// of n <= 64, m <= 72 constraints with sparse data (all F3), no LP block.
//
// Every thread of a work-group evaluates the control scalars from the same
// block reductions, so control flow is uniform and needs no broadcast.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/sdpa_batch.h"
#include "eig_dev.h"

namespace sdpa_amd {
namespace {

constexpr int NMAX = 64;        // block size limit
constexpr int MMAX = 72;        // constraint limit
constexpr int LDA = NMAX + 1;   // LDS leading dimension of block matrices
constexpr int LDB = MMAX + 1;   // of the Schur matrix
constexpr int kMats = 12;       // n x n work matrices per instance
constexpr int kLz = 64;         // Lanczos tridiagonal limit (sqrt(64) + 10 steps)

enum { noINFO, pFEAS, dFEAS, pdFEAS, pdINF, pFEAS_dINF, pINF_dFEAS, pdOPT, pUNBD, dUNBD };

struct BParams {
  int maxIteration;
  double epsilonStar, epsilonDash, betaStar, betaBar, gammaStar, omegaStar, lowerBound, upperBound;
};

struct BData {
  int count;
  const int* n;
  const int* m;
  const long* coff;
  const double* C;
  const long* boff;
  const double* b;
  const long* pptr;
  const int* cptr;
  const int* er;
  const int* ec;
  const double* ev;
  const double* lambda;
  const long* woff;  // work slab of each instance (kMats n^2 doubles)
  double* work;
  double* out;
};

template <int BT>
struct Shared {
  double A[NMAX * LDA];   // block-matrix scratch (GEMM operand, factor)
  double B2[NMAX * LDA];  // second operand / inverse
  double S[MMAX * LDB];   // Schur matrix and its factor
  double y[MMAX], dy[MMAX], g[MMAX], p[MMAX], bv[MMAX], tv[MMAX];
  double lq[NMAX], lqo[NMAX], lr[NMAX], lw[NMAX], dg[kLz], od[kLz];
  double red[BT / 64 + 1];
  double trisc[2 * (BT / 64)];
  double sc[4];
  int info;
};

template <int BT>
__device__ __forceinline__ double bsum(double v, Shared<BT>& sh) {
  v = wsum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < BT / 64; ++i) s += sh.red[i];
  return s;
}

template <int BT>
__device__ __forceinline__ double bmax(double v, Shared<BT>& sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < BT / 64; ++i) s = fmax(s, sh.red[i]);
  return s;
}

// sum of a[e] b[e] over len elements (global or LDS)
template <int BT>
__device__ double dot(const double* a, const double* b, int len, Shared<BT>& sh) {
  double s = 0.0;
  for (int e = threadIdx.x; e < len; e += BT) s += a[e] * b[e];
  return bsum(s, sh);
}

// out = a A + b B + c Cm (B, Cm may be null)
template <int BT>
__device__ void axpby(double* out, double a, const double* A, double b, const double* B, double c,
                      const double* Cm, int len) {
  for (int e = threadIdx.x; e < len; e += BT) {
    double v = a * A[e];
    if (B) v += b * B[e];
    if (Cm) v += c * Cm[e];
    out[e] = v;
  }
  __syncthreads();
}

struct Prob {
  int n, m;
  const double* C;
  const double* b;
  const int* cp;  // m + 1 global entry indices
  const int* er;
  const int* ec;
  const double* ev;
};

// out_i = A_i . D (symmetric inner product of the stored upper entries)
template <int BT>
__device__ void adot(const Prob& P, const double* D, double* out) {
  for (int i = threadIdx.x; i < P.m; i += BT) {
    double s = 0.0;
    for (int e = P.cp[i]; e < P.cp[i + 1]; ++e) {
      const int r = P.er[e], c = P.ec[e];
      double d = D[r + P.n * c];
      if (r != c) d += D[c + P.n * r];
      s += P.ev[e] * d;
    }
    out[i] = s;
  }
  __syncthreads();
}

// out = sum_i y_i A_i (entries scattered by one thread, in order)
template <int BT>
__device__ void sum_yA(const Prob& P, const double* y, double* out) {
  const int nn = P.n * P.n;
  for (int e = threadIdx.x; e < nn; e += BT) out[e] = 0.0;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < P.m; ++i)
      for (int e = P.cp[i]; e < P.cp[i + 1]; ++e) {
        const int r = P.er[e], c = P.ec[e];
        out[r + P.n * c] += y[i] * P.ev[e];
        if (r != c) out[c + P.n * r] += y[i] * P.ev[e];
      }
  }
  __syncthreads();
}

// C = op(A) op(B) (global n x n, ld n), through LDS
template <int BT>
__device__ void gemm(int n, const double* A, int ta, const double* B, int tb, double* C,
                     Shared<BT>& sh) {
  const int nn = n * n;
  for (int e = threadIdx.x; e < nn; e += BT) {
    const int i = e % n, j = e / n;
    sh.A[(ta ? j : i) + LDA * (ta ? i : j)] = A[e];
    sh.B2[(tb ? j : i) + LDA * (tb ? i : j)] = B[e];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nn; e += BT) {
    const int i = e % n, j = e / n;
    double s0 = 0.0, s1 = 0.0;
    int k = 0;
    for (; k + 1 < n; k += 2) {
      s0 += sh.A[i + LDA * k] * sh.B2[k + LDA * j];
      s1 += sh.A[i + LDA * (k + 1)] * sh.B2[(k + 1) + LDA * j];
    }
    if (k < n) s0 += sh.A[i + LDA * k] * sh.B2[k + LDA * j];
    C[e] = s0 + s1;
  }
  __syncthreads();
}

// in-place lower Cholesky of the n x n matrix M (LDS, ld) -- mode 0: LAPACK
// (a pivot that is not > 0 fails), mode 1: rATL_dpotrfL's rule (pivot < -1e-6
// fails, pivot < 1e-14 becomes 1e100).  Returns 0 or the 1-based column.
template <int BT>
__device__ int chol(double* M, int ld, int n, int mode, Shared<BT>& sh) {
  if (threadIdx.x == 0) sh.info = 0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (threadIdx.x == 0) {
      double d = M[k + ld * k];
      bool bad;
      if (mode == 0) {
        bad = !(d > 0.0);
      } else {
        bad = d < -1.0e-6;
        if (!bad && d < 1.0e-14) d = 1.0e100;
      }
      if (bad) sh.info = k + 1;
      M[k + ld * k] = bad ? 1.0 : sqrt(d);
    }
    __syncthreads();
    if (sh.info) return sh.info;
    const double piv = M[k + ld * k];
    for (int i = k + 1 + threadIdx.x; i < n; i += BT) M[i + ld * k] /= piv;
    __syncthreads();
    const int rem = n - k - 1;
    for (int e = threadIdx.x; e < rem * rem; e += BT) {
      const int i = k + 1 + e % rem, j = k + 1 + e / rem;
      if (i >= j) M[i + ld * j] -= M[i + ld * k] * M[j + ld * k];
    }
    __syncthreads();
  }
  return 0;
}

// W = L^-1 (LDS, lower, zero above) from the lower factor L (LDS)
template <int BT>
__device__ void trinv(const double* L, double* W, int n) {
  const int j = threadIdx.x;
  if (j < n) {
    for (int i = 0; i < j; ++i) W[i + LDA * j] = 0.0;
    for (int i = j; i < n; ++i) {
      double s = (i == j) ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) s -= L[i + LDA * k] * W[k + LDA * j];
      W[i + LDA * j] = s / L[i + LDA * i];
    }
  }
  __syncthreads();
}

// x := L^-T L^-1 x with the Schur factor in sh.S (x in LDS)
template <int BT>
__device__ void schur_solve(int m, double* x, Shared<BT>& sh) {
  for (int k = 0; k < m; ++k) {
    if (threadIdx.x == 0) x[k] /= sh.S[k + LDB * k];
    __syncthreads();
    for (int i = k + 1 + threadIdx.x; i < m; i += BT) x[i] -= sh.S[i + LDB * k] * x[k];
    __syncthreads();
  }
  for (int k = m - 1; k >= 0; --k) {
    if (threadIdx.x == 0) {
      double s = x[k];
      for (int i = k + 1; i < m; ++i) s -= sh.S[i + LDB * k] * x[i];
      x[k] = s / sh.S[k + LDB * k];
    }
    __syncthreads();
  }
}

// calF3 with A_j the owner (outer loop), sdpa_newton.cpp:1319-1403
__device__ double calF3(const Prob& P, const double* X, const double* Zi, int i, int j) {
  const int n = P.n;
  double ret = 0.0;
  for (int e1 = P.cp[j]; e1 < P.cp[j + 1]; ++e1) {
    const int a = P.er[e1], b = P.ec[e1];
    const double* Xa = X + n * a;
    const double* Xb = X + n * b;
    const double* Za = Zi + n * a;
    const double* Zb = Zi + n * b;
    double sum = 0.0;
    for (int e2 = P.cp[i]; e2 < P.cp[i + 1]; ++e2) {
      const int g = P.er[e2], d = P.ec[e2];
      double plu;
      if (a != b) {
        plu = (g != d) ? Zb[d] * Xa[g] + Zb[g] * Xa[d] + Za[d] * Xb[g] + Za[g] * Xb[d]
                       : Zb[d] * Xa[g] + Za[d] * Xb[g];
      } else {
        plu = (g != d) ? Zb[d] * Xa[g] + Zb[g] * Xa[d] : Zb[d] * Xa[g];
      }
      sum += P.ev[e2] * plu;
    }
    ret += P.ev[e1] * sum;
  }
  return ret;
}

// NonZeroEffect: n^2 for a dense A_i (>= 20% of the entries stored,
// sdpa_struct.cpp:518), else the entry count with off-diagonals twice
__device__ __forceinline__ int nzeff(const Prob& P, int i) {
  const int cnt = P.cp[i + 1] - P.cp[i];
  if ((double)cnt >= (double)P.n * P.n * 0.20) return P.n * P.n;
  int s = 0;
  for (int e = P.cp[i]; e < P.cp[i + 1]; ++e) s += (P.er[e] == P.ec[e]) ? 1 : 2;
  return s;
}

// the Schur matrix into sh.S (both triangles), F3 with SDPA's ownership rule
template <int BT>
__device__ void make_bmat(const Prob& P, const double* X, const double* Zi, Shared<BT>& sh) {
  const int m = P.m, np = m * (m + 1) / 2;
  for (int t = threadIdx.x; t < np; t += BT) {
    // pair (i, j), i >= j, column by column
    int j = 0, r = t;
    while (r >= m - j) { r -= m - j; ++j; }
    const int i = j + r;
    const int jnz = nzeff(P, j), inz = nzeff(P, i);
    // the owner is the one with the larger NonZeroEffect (ties: the smaller index)
    const bool jown = !((jnz < inz) || (inz == jnz && i < j));
    const double v = jown ? calF3(P, X, Zi, i, j) : calF3(P, X, Zi, j, i);
    sh.S[i + LDB * j] = v;
    sh.S[j + LDB * i] = v;
  }
  __syncthreads();
}

// smallest eigenvalue of the symmetric n x n matrix in sh.A by cyclic Jacobi
// (the dsyev branch of Lal::getMinEigen for n <= 20, sdpa_linear.cpp:145-182):
// thread k rotates row/column k; every thread takes the same decisions.
template <int BT>
__device__ double jacobi_min(int n, Shared<BT>& sh) {
  const int t = threadIdx.x;
  double* A = sh.A;
  for (int sweep = 0; sweep < 100; ++sweep) {
    __syncthreads();  // the last rotation's writes before the norms
    double off = 0.0, dv = 0.0;
    if (t < n) {
      for (int j = 0; j < t; ++j) off += A[t + LDA * j] * A[t + LDA * j];
      dv = A[t + LDA * t] * A[t + LDA * t];
    }
    off = bsum(off, sh);
    dv = bsum(dv, sh);
    // converged at an eps-relative off-diagonal norm
    if (off == 0.0 || off <= 1e-31 * dv) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        __syncthreads();
        const double apq = A[p + LDA * q];
        if (apq == 0.0) continue;
        const double app = A[p + LDA * p], aqq = A[q + LDA * q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(tt * tt + 1.0), s = tt * c;
        __syncthreads();
        if (t < n) {
          const double akp = A[t + LDA * p], akq = A[t + LDA * q];
          A[t + LDA * p] = c * akp - s * akq;
          A[t + LDA * q] = s * akp + c * akq;
        }
        __syncthreads();
        if (t < n) {
          const double apk = A[p + LDA * t], aqk = A[q + LDA * t];
          A[p + LDA * t] = c * apk - s * aqk;
          A[q + LDA * t] = s * apk + c * aqk;
        }
      }
  }
  __syncthreads();
  double mn = 1.0e300;
  for (int i = 0; i < n; ++i) mn = fmin(mn, A[i + LDA * i]);
  __syncthreads();
  return mn;
}

// Lal::getMinEigen on the symmetric M (global, n x n): the truncated Lanczos
// (n > 20) or all eigenvalues by Jacobi (n <= 20).
template <int BT>
__device__ double min_eig(const double* M, int n, Shared<BT>& sh) {
  const int tid = threadIdx.x;
  if (n <= 20) {
    for (int e = tid; e < n * n; e += BT) sh.A[(e % n) + LDA * (e / n)] = M[e];
    __syncthreads();
    return jacobi_min(n, sh);
  }
  double mn = 1.0e+51, mn_old = 1.0e+52, error = 1.0e+10;
  int k = 0, kk = 0;
  for (int i = tid; i < n; i += BT) { sh.lq[i] = 0.0; sh.lr[i] = 1.0; }
  double beta = sqrt((double)n);
  const double sqn = sqrt((double)n);
  while (k < n && k < sqn + 10 && k < kLz && beta > 1.0e-16 &&
         (fabs(mn - mn_old) > (1.0e-5) * fabs(mn) + (1.0e-8) ||
          fabs(error * beta) > (1.0e-2) * fabs(mn) + (1.0e-4))) {
    const double inv = 1.0 / beta;
    __syncthreads();
    for (int i = tid; i < n; i += BT) { sh.lqo[i] = sh.lq[i]; sh.lq[i] = sh.lr[i] * inv; }
    __syncthreads();
    for (int i = tid; i < n; i += BT) {
      double a = 0.0;
      for (int j = 0; j < n; ++j) a += M[i + n * j] * sh.lq[j];
      sh.lw[i] = a;
    }
    __syncthreads();
    double a = 0.0;
    for (int i = tid; i < n; i += BT) a += sh.lq[i] * sh.lw[i];
    const double alpha = bsum(a, sh);
    if (tid == 0) sh.dg[k] = alpha;
    for (int i = tid; i < n; i += BT) sh.lr[i] = (sh.lw[i] - sh.lq[i] * alpha) - sh.lqo[i] * beta;
    if (kk >= sqrt((double)k) || k == n - 1 || k > sqrt((double)n + 9)) {
      kk = 0;
      __syncthreads();
      double th = 0.0, zl = 0.0;
      tridiag_min_block(k + 1, sh.dg, sh.od, th, zl, BT, sh.trisc);
      if (tid == 0) { sh.sc[0] = th; sh.sc[1] = zl; }
      __syncthreads();
      mn_old = mn;
      mn = sh.sc[0];
      error = sh.sc[1];
    }
    double v2 = 0.0;
    for (int i = tid; i < n; i += BT) v2 += sh.lr[i] * sh.lr[i];
    beta = sqrt(bsum(v2, sh));
    if (tid == 0) sh.od[k] = beta;
    ++k;
    ++kk;
  }
  __syncthreads();
  return mn - fabs(error * beta);
}

// invChol (LXi, LZi) and Zi from X, Z; 0 or 1 (a factorization failed)
template <int BT>
__device__ int inv_chol(int n, const double* X, const double* Z, double* LXi, double* LZi,
                        double* Zi, Shared<BT>& sh) {
  for (int w = 0; w < 2; ++w) {
    const double* S = w ? Z : X;
    for (int e = threadIdx.x; e < n * n; e += BT) sh.A[(e % n) + LDA * (e / n)] = S[e];
    __syncthreads();
    if (chol(sh.A, LDA, n, 0, sh)) return 1;
    for (int e = threadIdx.x; e < n * n; e += BT) {  // zero the upper part
      const int i = e % n, j = e / n;
      if (i < j) sh.A[i + LDA * j] = 0.0;
    }
    __syncthreads();
    trinv<BT>(sh.A, sh.B2, n);
    double* out = w ? LZi : LXi;
    for (int e = threadIdx.x; e < n * n; e += BT) out[e] = sh.B2[(e % n) + LDA * (e / n)];
    __syncthreads();
  }
  gemm(n, LZi, 1, LZi, 0, Zi, sh);  // Zi = L^-T L^-1
  return 0;
}

template <int BT>
__global__ __launch_bounds__(BT) void k_batch_solve(BData Dt, BParams Pm) {
  __shared__ Shared<BT> sh;
  const int tid = threadIdx.x;
  for (int k = blockIdx.x; k < Dt.count; k += gridDim.x) {
    Prob P;
    P.n = Dt.n[k];
    P.m = Dt.m[k];
    P.C = Dt.C + Dt.coff[k];
    P.b = Dt.b + Dt.boff[k];
    P.cp = Dt.cptr + Dt.pptr[k];
    P.er = Dt.er;
    P.ec = Dt.ec;
    P.ev = Dt.ev;
    const int n = P.n, m = P.m, nn = n * n;
    double* wk = Dt.work + Dt.woff[k];
    double* X = wk;
    double* Z = wk + nn;
    double* Zi = wk + 2 * nn;
    double* LXi = wk + 3 * nn;
    double* LZi = wk + 4 * nn;
    double* Dx = wk + 5 * nn;
    double* Dz = wk + 6 * nn;
    double* R = wk + 7 * nn;
    double* Dm = wk + 8 * nn;
    double* W = wk + 9 * nn;
    double* T = wk + 10 * nn;
    double* XDZi = wk + 11 * nn;
    const double lam = Dt.lambda[k];
    const double nDim = n;
    // identity start (Solutions::initialize, sdpa_dataset.cpp:47-74)
    for (int e = tid; e < nn; e += BT) {
      const bool d = (e % n) == (e / n);
      X[e] = d ? lam : 0.0;
      Z[e] = d ? lam : 0.0;
      LXi[e] = d ? 1.0 / sqrt(lam) : 0.0;
      LZi[e] = d ? 1.0 / sqrt(lam) : 0.0;
      Zi[e] = d ? 1.0 / lam : 0.0;
    }
    for (int i = tid; i < m; i += BT) { sh.y[i] = 0.0; sh.bv[i] = P.b[i]; }
    __syncthreads();
    const double muInit = lam * lam;
    double muCur = muInit;
    // residuals: p = b - A(X), D = C - sum y A - Z
    auto residuals = [&](double& nP, double& nD) {
      adot<BT>(P, X, sh.p);
      for (int i = tid; i < m; i += BT) sh.p[i] = sh.bv[i] - sh.p[i];
      __syncthreads();
      sum_yA<BT>(P, sh.y, Dm);
      axpby<BT>(Dm, -1.0, Dm, 1.0, P.C, -1.0, Z, nn);
      double a = 0.0, c = 0.0;
      for (int i = tid; i < m; i += BT) a = fmax(a, fabs(sh.p[i]));
      for (int e = tid; e < nn; e += BT) c = fmax(c, fabs(Dm[e]));
      nP = bmax(a, sh);
      nD = bmax(c, sh);
    };
    double normP, normD;
    residuals(normP, normD);
    const double initNormP = normP, initNormD = normD;
    double betaVal = Pm.betaStar;
    double thetaP = (initNormP < Pm.epsilonDash) ? 0.0 : 1.0;
    double thetaD = (initNormD < Pm.epsilonDash) ? 0.0 : 1.0;
    double rho = 1.0;
    double etaP = Pm.omegaStar * nDim * muInit, etaD = etaP;
    double objP = dot(P.C, X, nn, sh), objD = dot(sh.bv, sh.y, m, sh);
    int phase = noINFO, failure = 0;
    double alphaP = 0.0, alphaD = 0.0;
    // Phase::updateCheck (sdpa_parts.cpp:1150-1242)
    auto update_phase = [&]() -> bool {
      const double NONZERO = 1.0e-6, acc = Pm.epsilonDash;
      int v = noINFO;
      if (normP <= acc) v = (normD <= acc) ? pdFEAS : pFEAS;
      if (v == noINFO && normD <= acc) v = dFEAS;
      phase = v;
      if (v == pdFEAS) {
        const double mean = (fabs(objP) + fabs(objD)) / 2.0;
        const double gap = fabs(objP - objD);
        const double dom = mean < 1.0 ? 1.0 : mean;
        if (gap / dom <= Pm.epsilonStar) { phase = pdOPT; return false; }
      }
      if (v == noINFO && rho > 1.0 + NONZERO) { phase = pdINF; return false; }
      if (v == pFEAS) {
        if (objP <= -Pm.upperBound) { phase = pUNBD; return false; }
        if (rho > 1.0 + NONZERO) { phase = pFEAS_dINF; return false; }
      }
      if (v == dFEAS) {
        if (objD >= -Pm.lowerBound) { phase = dUNBD; return false; }
        if (rho > 1.0 + NONZERO) { phase = pINF_dFEAS; return false; }
      }
      return true;
    };
    update_phase();
    // r_zinv = target Zinv - X (- Dx Dz Zinv), minus X D Zinv when infeasible;
    // g = -A(.) + p   (Newton::Make_gVec)
    auto make_gvec = [&](int corrector, double target, int infeasible) {
      axpby<BT>(R, target, Zi, -1.0, X, 0.0, nullptr, nn);
      if (corrector) {
        gemm(n, Dx, 0, Dz, 0, T, sh);
        gemm(n, T, 0, Zi, 0, W, sh);
        axpby<BT>(R, 1.0, R, -1.0, W, 0.0, nullptr, nn);
      }
      const double* Dd = R;
      if (infeasible) {
        if (!corrector) {
          gemm(n, X, 0, Dm, 0, T, sh);
          gemm(n, T, 0, Zi, 0, XDZi, sh);
        }
        axpby<BT>(W, 1.0, R, -1.0, XDZi, 0.0, nullptr, nn);
        Dd = W;
      }
      adot<BT>(P, Dd, sh.g);
      for (int i = tid; i < m; i += BT) sh.g[i] = -sh.g[i] + sh.p[i];
      __syncthreads();
    };
    // Dz = -sum dy A (+ D); Dx = sym(r_zinv - X Dz Zinv)
    auto make_direction = [&](int infeasible) {
      sum_yA<BT>(P, sh.dy, Dz);
      axpby<BT>(Dz, -1.0, Dz, infeasible ? 1.0 : 0.0, infeasible ? Dm : nullptr, 0.0, nullptr, nn);
      gemm(n, X, 0, Dz, 0, T, sh);
      gemm(n, T, 0, Zi, 0, W, sh);
      axpby<BT>(Dx, 1.0, R, -1.0, W, 0.0, nullptr, nn);
      for (int e = tid; e < nn; e += BT) {
        const int i = e % n, j = e / n;
        if (i > j) {
          const double s = (Dx[i + n * j] + Dx[j + n * i]) * 0.5;
          Dx[i + n * j] = s;
          Dx[j + n * i] = s;
        }
      }
      __syncthreads();
    };
    int it = 0;
    while (update_phase() && it < Pm.maxIteration) {
      const bool affine = !(phase == noINFO || phase == pFEAS || phase == dFEAS);
      if (phase == pdFEAS) betaVal = 0.0;
      else betaVal = affine ? 2.0 : Pm.betaBar;
      const int infeasible = (phase == pFEAS || phase == noINFO) ? 1 : 0;
      // Newton::Mehrotra(PREDICTOR)
      make_gvec(0, betaVal * muCur, infeasible);
      make_bmat(P, X, Zi, sh);
      if (chol(sh.S, LDB, m, 1, sh)) { failure = 1; break; }
      for (int i = tid; i < m; i += BT) sh.dy[i] = sh.g[i];
      __syncthreads();
      schur_solve(m, sh.dy, sh);
      make_direction(infeasible);
      double q[7];
      q[0] = dot(P.C, Dx, nn, sh);
      q[1] = dot(sh.bv, sh.dy, m, sh);
      q[2] = dot(X, Dz, nn, sh);
      q[3] = dot(Dx, Z, nn, sh);
      q[4] = dot(Dx, Dz, nn, sh);
      alphaP = alphaD = 0.9;
      if (phase == noINFO || phase == dFEAS) {
        if (alphaP > 1.0) alphaP = 1.0;
      } else if (q[0] > 0.0) {
        if (alphaP > alphaD) alphaP = alphaD;
        if (alphaP > 1.0) alphaP = 1.0;
      }
      if (phase == noINFO || phase == pFEAS) {
        if (alphaD > 1.0) alphaD = 1.0;
      } else if (q[1] < 0.0) {
        if (alphaD > alphaP) alphaD = alphaP;
        if (alphaD > 1.0) alphaD = 1.0;
      }
      {
        const double muT = muCur + (alphaP * q[3] + alphaD * q[2] + alphaP * alphaD * q[4]) / nDim;
        double v = muT / muCur;
        if (v < 1.0) v = v * v;
        if (phase == pdFEAS) {
          if (v < Pm.betaStar) v = Pm.betaStar;
          if (v > 1.0) v = 1.0;
        } else if (v < Pm.betaBar) {
          v = Pm.betaBar;
        }
        betaVal = v;
      }
      // Newton::Mehrotra(CORRECTOR)
      make_gvec(1, betaVal * muCur, infeasible);
      for (int i = tid; i < m; i += BT) sh.dy[i] = sh.g[i];
      __syncthreads();
      schur_solve(m, sh.dy, sh);
      make_direction(infeasible);
      // StepLength::MehrotraCorrector
      gemm(n, Dx, 0, LXi, 1, T, sh);
      gemm(n, LXi, 0, T, 0, W, sh);
      const double lamX = min_eig(W, n, sh);
      gemm(n, Dz, 0, LZi, 1, T, sh);
      gemm(n, LZi, 0, T, 0, W, sh);
      const double lamZ = min_eig(W, n, sh);
      const double alphaBD = 100.0;
      alphaP = (-lamX > 1.0 / alphaBD) ? -1.0 / lamX : alphaBD;
      alphaD = (-lamZ > 1.0 / alphaBD) ? -1.0 / lamZ : alphaBD;
      alphaP *= Pm.gammaStar;
      alphaD *= Pm.gammaStar;
      q[0] = dot(P.C, Dx, nn, sh);
      q[1] = dot(sh.bv, sh.dy, m, sh);
      q[2] = dot(X, Dz, nn, sh);
      q[3] = dot(Dx, Z, nn, sh);
      q[4] = dot(Dx, Dz, nn, sh);
      q[5] = dot(sh.bv, sh.y, m, sh);
      q[6] = dot(P.C, X, nn, sh);
      if (phase == noINFO || phase == dFEAS) {
        if (alphaP > 1.0) alphaP = 1.0;
      } else if (q[0] > 0.0) {
        if (alphaP > alphaD) alphaP = alphaD;
        if (alphaP > 1.0) alphaP = 1.0;
      }
      if (phase == noINFO || phase == pFEAS) {
        if (alphaD > 1.0) alphaD = 1.0;
      } else if (q[1] < 0.0) {
        if (alphaD > alphaP) alphaD = alphaP;
        if (alphaD > 1.0) alphaD = 1.0;
      }
      if (!affine && (phase == noINFO || phase == pFEAS || phase == dFEAS)) {
        const double xi = 3.0;
        double thetaMax = fmax((1.0 - alphaP) * thetaP, (1.0 - alphaD) * thetaD);
        double muNew = muCur + (alphaP * q[3] + alphaD * q[2] + alphaP * alphaD * q[4]) / nDim;
        while (thetaMax * thetaMax * muInit > xi * muNew) {
          const double aMax = 0.95 * fmax(alphaP, alphaD);
          alphaP = fmin(alphaP, aMax);
          alphaD = fmin(alphaD, aMax);
          thetaMax = fmax((1.0 - alphaP) * thetaP, (1.0 - alphaD) * thetaD);
          muNew = muCur + (alphaP * q[3] + alphaD * q[2] + alphaP * alphaD * q[4]) / nDim;
          if (alphaP < 1.0e-6 && alphaD < 1.0e-6) break;
        }
      }
      if (phase == pdFEAS) {
        const double incD = q[1] * alphaD, incP = q[0] * alphaP;
        const double maxRatio = (q[5] - q[6]) / (incP - incD);
        if (maxRatio > 0.0 && maxRatio < 1.0) { alphaP *= maxRatio; alphaD *= maxRatio; }
      }
      // Solutions::update (sdpa_dataset.cpp:145-171)
      axpby<BT>(X, 1.0, X, alphaP, Dx, 0.0, nullptr, nn);
      for (int i = tid; i < m; i += BT) sh.y[i] += alphaD * sh.dy[i];
      axpby<BT>(Z, 1.0, Z, alphaD, Dz, 0.0, nullptr, nn);
      if (alphaP < 1.0e-4 && alphaD < 1.0e-4) { failure = 2; break; }
      if (inv_chol(n, X, Z, LXi, LZi, Zi, sh)) { failure = 3; break; }
      if (!affine) {
        thetaP = fabs((1.0 - alphaP) * thetaP);
        thetaD = fabs((1.0 - alphaD) * thetaD);
      }
      double trZ = 0.0, trX = 0.0;
      for (int i = tid; i < n; i += BT) { trZ += Z[i + n * i]; trX += X[i + n * i]; }
      trZ = bsum(trZ, sh);
      trX = bsum(trX, sh);
      const double xz = dot(X, Z, nn, sh);
      objP = dot(P.C, X, nn, sh);
      objD = dot(sh.bv, sh.y, m, sh);
      muCur = xz / nDim;
      residuals(normP, normD);
      thetaP = (initNormP > Pm.epsilonDash * 1.0e-2) ? normP / initNormP : 0.0;
      thetaD = (initNormD > Pm.epsilonDash * 1.0e-2) ? normD / initNormD : 0.0;
      // SolveInfo::update (sdpa_parts.cpp:914-1060)
      rho = 0.0;
      const double x0z0 = nDim * muInit, xzn = nDim * muCur;
      const double x0z = lam * trZ, xz0 = lam * trX;
      const double acc = Pm.epsilonDash, omega = Pm.omegaStar;
      if (normP <= acc && xz0 < etaP) etaP = xz0;
      if (normD <= acc && x0z < etaD) etaD = x0z;
      if (normP > acc && normD <= acc)
        rho = thetaP * x0z / ((thetaP + (1.0 - thetaP) * omega) * etaD + xzn);
      if (normP <= acc && normD > acc)
        rho = thetaD * xz0 / ((thetaD + (1.0 - thetaD) * omega) * etaP + xzn);
      if (normP > acc && normD > acc)
        rho = (thetaD * xz0 + thetaP * x0z) /
              ((thetaP * thetaD + omega * (thetaP * (1.0 - thetaD) + (1.0 - thetaP) * thetaD)) *
                   x0z0 + xzn);
      ++it;
    }
    residuals(normP, normD);
    if (tid == 0) {
      double* o = Dt.out + 8 * (size_t)k;
      o[0] = phase;
      o[1] = it;
      o[2] = objP;
      o[3] = objD;
      o[4] = normP;
      o[5] = normD;
      o[6] = muCur;
      o[7] = failure;
    }
    __syncthreads();
  }
}

}  // namespace
}  // namespace sdpa_amd

using namespace sdpa_amd;

#define BCHK(x)                          \
  do {                                   \
    if ((x) != hipSuccess) return -100;  \
  } while (0)

extern "C" int sdpa_batch_solve(int device, int count, const int* n, const int* m,
                                const long* coff, const double* C, long nC, const long* boff,
                                const double* b, long nb, const long* pptr, const int* cptr,
                                long ncptr, const int* er, const int* ec, const double* ev,
                                long nent, const double* lambda, const double* param,
                                int threads, double* out, double* kernel_ms) {
  if (count <= 0 || !n || !m || !coff || !C || !boff || !b || !pptr || !cptr || !er || !ec ||
      !ev || !lambda || !param || !out)
    return -1;
  if (threads != 128 && threads != 256) return -2;
  std::vector<long> woff(count);
  long wtot = 0;
  for (int k = 0; k < count; ++k) {
    if (n[k] < 1 || n[k] > NMAX || m[k] < 1 || m[k] > MMAX) return -3;
    if (coff[k] < 0 || coff[k] + (long)n[k] * n[k] > nC || boff[k] < 0 || boff[k] + m[k] > nb ||
        pptr[k] < 0 || pptr[k] + m[k] + 1 > ncptr)
      return -4;
    for (int i = 0; i <= m[k]; ++i) {
      const int e = cptr[pptr[k] + i];
      if (e < 0 || e > nent || (i > 0 && e < cptr[pptr[k] + i - 1])) return -5;
    }
    woff[k] = wtot;
    wtot += (long)kMats * n[k] * n[k];
  }
  for (long e = 0; e < nent; ++e)
    if (er[e] < 0 || ec[e] < er[e]) return -6;
  for (int k = 0; k < count; ++k)
    for (int e = cptr[pptr[k]]; e < cptr[pptr[k] + m[k]]; ++e)
      if (ec[e] >= n[k]) return -6;
  BCHK(hipSetDevice(device));
  hipStream_t s;
  BCHK(hipStreamCreate(&s));
  int *dn = nullptr, *dm = nullptr, *dcp = nullptr, *der = nullptr, *dec = nullptr;
  long *dco = nullptr, *dbo = nullptr, *dpp = nullptr, *dwo = nullptr;
  double *dC = nullptr, *db = nullptr, *dev = nullptr, *dl = nullptr, *dw = nullptr, *dout = nullptr;
  int rc = 0;
  auto up = [&](void** dst, const void* src, size_t bytes) -> int {
    if (hipMalloc(dst, bytes ? bytes : 8) != hipSuccess) return -100;
    if (bytes && hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
      return -100;
    return 0;
  };
  do {
    if ((rc = up((void**)&dn, n, sizeof(int) * count))) break;
    if ((rc = up((void**)&dm, m, sizeof(int) * count))) break;
    if ((rc = up((void**)&dco, coff, sizeof(long) * count))) break;
    if ((rc = up((void**)&dC, C, sizeof(double) * nC))) break;
    if ((rc = up((void**)&dbo, boff, sizeof(long) * count))) break;
    if ((rc = up((void**)&db, b, sizeof(double) * nb))) break;
    if ((rc = up((void**)&dpp, pptr, sizeof(long) * count))) break;
    if ((rc = up((void**)&dcp, cptr, sizeof(int) * ncptr))) break;
    if ((rc = up((void**)&der, er, sizeof(int) * nent))) break;
    if ((rc = up((void**)&dec, ec, sizeof(int) * nent))) break;
    if ((rc = up((void**)&dev, ev, sizeof(double) * nent))) break;
    if ((rc = up((void**)&dl, lambda, sizeof(double) * count))) break;
    if ((rc = up((void**)&dwo, woff.data(), sizeof(long) * count))) break;
    if (hipMalloc(&dw, sizeof(double) * (wtot ? wtot : 1)) != hipSuccess) { rc = -100; break; }
    if (hipMalloc(&dout, sizeof(double) * 8 * count) != hipSuccess) { rc = -100; break; }
    BData D{count, dn, dm, dco, dC, dbo, db, dpp, dcp, der, dec, dev, dl, dwo, dw, dout};
    BParams P{(int)param[0], param[1], param[2], param[3], param[4], param[5], param[6], param[7],
              param[8]};
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
      rc = -100;
      break;
    }
    // persistent work-groups: one per CU (LDS of ~110 KB allows one), each
    // looping over instances
    const int grid = count < ncu ? count : ncu;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { rc = -100; break; }
    (void)hipEventRecord(e0, s);
    if (threads == 256)
      hipLaunchKernelGGL(k_batch_solve<256>, dim3(grid), dim3(256), 0, s, D, P);
    else
      hipLaunchKernelGGL(k_batch_solve<128>, dim3(grid), dim3(128), 0, s, D, P);
    (void)hipEventRecord(e1, s);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      rc = -101;
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (kernel_ms) *kernel_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (hipMemcpy(out, dout, sizeof(double) * 8 * count, hipMemcpyDeviceToHost) != hipSuccess)
      rc = -100;
  } while (0);
  for (void* p : {(void*)dn, (void*)dm, (void*)dcp, (void*)der, (void*)dec, (void*)dco, (void*)dbo,
                  (void*)dpp, (void*)dwo, (void*)dC, (void*)db, (void*)dev, (void*)dl, (void*)dw,
                  (void*)dout})
    if (p) (void)hipFree(p);
  (void)hipStreamDestroy(s);
  return rc;
}
