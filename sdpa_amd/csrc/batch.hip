// batch.hip -- many small SDPs solved by one persistent launch.  One
// work-group per instance runs SDPA's whole Mehrotra predictor-corrector
// loop (SDPA::solve, sdpa_solve.cpp:59-233): the Newton direction (Make_gVec
// sdpa_newton.cpp:817-883, F3 bMat 1319-1403, rATL_dpotrfL sdpa_dpotrf.cpp:
// 218-265, compute_DzMat/DxMat 2718-2752), the step lengths
// (Jal::getMinEigen, sdpa_jordan.cpp:55-90: Lanczos for n > 20, all
// eigenvalues for n <= 20) and the control scalars (sdpa_parts.cpp:
// 493-1242) -- the formulas of the single-problem path, with every scalar
// decision taken on the device.  Instances: one SDP block of n <= 64,
// m <= 72 constraints with at most kEnt stored entries, no LP block.
//
// Layout per work-group: the instance's constraint entries, Schur matrix and
// GEMM operands live in LDS (about 137 KB, one work-group per CU); the n x n
// iterates and directions live in a per-instance global slab (L2-resident).
// Work-groups take instances from a queue, largest first.
//
// Every thread evaluates the control scalars from the same block
// reductions, so control flow is uniform and needs no broadcast.  The
// sequential kernels of one iteration use one wave where a chain of
// dependent steps would otherwise pay a barrier per step: the Cholesky
// factors and inverses of X and Z (one wave each, concurrently), the
// triangular solves with the Schur factor, and the Householder reduction of
// small blocks.  Products of n x n matrices are FP64 MFMA tiles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "../../include/sdpa_batch.h"
#include "eig_dev.h"

namespace sdpa_amd {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NMAX = 64;        // block size limit
constexpr int MMAX = 72;        // constraint limit
constexpr int kEnt = 1536;      // stored constraint entries per instance
constexpr int LDA = NMAX + 1;   // LDS leading dimension of block matrices
constexpr int LDB = MMAX + 1;   // of the Schur matrix
constexpr int kMats = 12;       // n x n work matrices per instance
constexpr int kLz = 64;         // Lanczos tridiagonal limit (sqrt(64) + 10 steps)
constexpr int kPadCols = 16;    // scratch columns behind A and B2 (straight-line loops)
constexpr int kDefaultHeavy = 64;  // F3 pairs with more entry products: one wave each

enum { noINFO, pFEAS, dFEAS, pdFEAS, pdINF, pFEAS_dINF, pINF_dFEAS, pdOPT, pUNBD, dUNBD };

struct BParams {
  int maxIteration;
  double epsilonStar, epsilonDash, betaStar, betaBar, gammaStar, omegaStar, lowerBound, upperBound;
  // launch configuration (sdpa_batch_config): F3 pairs with more entry
  // products than `heavy` take one wave each; precision 1 factors the
  // equilibrated Schur matrix in FP32 and refines the solves in FP64 while
  // mu >= muSwitch
  int heavy;
  int precision;
  double muSwitch;
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
  const long* woff;   // work slab of each instance (kMats n^2 doubles), or null:
                      // one slab of kMats NMAX^2 doubles per work-group
  const int* order;   // instances, largest first
  unsigned* next;     // queue head (zero at launch)
  double* work;
  double* out;
  double* stats;      // 4 per instance (may be null): FP32 factorizations, FP64
                      // factorizations, refinement steps, refinement stalls
};

template <int BT>
struct Shared {
  double A[(NMAX + kPadCols) * LDA];   // GEMM operand / X factor / step-length matrix
  double B2[(NMAX + kPadCols) * LDA];  // GEMM operand / Z factor
  double S[MMAX * LDB];   // Schur matrix and its factor (lower)
  double ev[kEnt];        // constraint entries: value, row, column, constraint
  double y[MMAX], dy[MMAX], g[MMAX], p[MMAX], bv[MMAX], piv[MMAX];
  double lq[NMAX], lqo[NMAX], lr[NMAX], lw[NMAX], dg[kLz], od[kLz];
  double red[8 * (BT / 64)];
  double trisc[2 * (BT / 64)];
  double sc[4];
  double pv[2][NMAX];         // pivots of the X and Z factorizations
  double eqs[MMAX];           // precision 1: equilibration s_i = B_ii^-1/2
  double xr[MMAX], rr[MMAX];  // precision 1: refined solution and residual
  float fd[MMAX];             // precision 1: pivots of the FP32 factor
  unsigned short perm[kEnt];  // entries in position-major order
  unsigned short cp[MMAX + 1];
  unsigned char er[kEnt], ec[kEnt], con[kEnt], head[kEnt];
  int nz[MMAX];
  int info[2];
  int task;
};

// the same value in every lane, as the compiler's uniform (scalar) value
__device__ __forceinline__ double uni(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// block sums of K values; every thread receives all K
template <int BT, int K>
__device__ __forceinline__ void bsums(double (&v)[K], Shared<BT>& sh) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wsum(v[k]);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh.red[(threadIdx.x >> 6) * K + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int w = 0; w < BT / 64; ++w) s += sh.red[w * K + k];
    v[k] = s;
  }
}

template <int BT>
__device__ __forceinline__ double bsum(double v, Shared<BT>& sh) {
  double a[1] = {v};
  bsums<BT, 1>(a, sh);
  return a[0];
}

template <int BT>
__device__ __forceinline__ void bmax2(double& a, double& b, Shared<BT>& sh) {
  a = wmax(a);
  b = wmax(b);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    sh.red[2 * (threadIdx.x >> 6)] = a;
    sh.red[2 * (threadIdx.x >> 6) + 1] = b;
  }
  __syncthreads();
  a = b = 0.0;
  for (int w = 0; w < BT / 64; ++w) {
    a = fmax(a, sh.red[2 * w]);
    b = fmax(b, sh.red[2 * w + 1]);
  }
}

// out = a A + b B + c Cm (B, Cm may be null)
template <int BT>
__device__ __forceinline__ void axpby(double* out, double a, const double* A, double b, const double* B, double c,
                      const double* Cm, int len) {
  for (int e = threadIdx.x; e < len; e += BT) {
    double v = a * A[e];
    if (B) v += b * B[e];
    if (Cm) v += c * Cm[e];
    out[e] = v;
  }
  __syncthreads();
}

// The instance's constraint data into LDS: entries, per-constraint pointers
// and NonZeroEffect (n^2 for a dense A_i, >= 20% of the entries stored,
// sdpa_struct.cpp:518; else the entry count with off-diagonals twice), and
// the entries in position-major order (position r + n c, then constraint)
// with run heads, for the deterministic sum_i y_i A_i.
template <int BT>
__device__ __forceinline__ int stage(const BData& Dt, int k, int n, int m, Shared<BT>& sh) {
  const int tid = threadIdx.x;
  const int* gcp = Dt.cptr + Dt.pptr[k];
  const int e0 = gcp[0];
  const int ne = gcp[m] - e0;
  for (int i = tid; i <= m; i += BT) sh.cp[i] = (unsigned short)(gcp[i] - e0);
  for (int e = tid; e < ne; e += BT) {
    sh.er[e] = (unsigned char)Dt.er[e0 + e];
    sh.ec[e] = (unsigned char)Dt.ec[e0 + e];
    sh.ev[e] = Dt.ev[e0 + e];
  }
  const double* b = Dt.b + Dt.boff[k];
  for (int i = tid; i < m; i += BT) {
    sh.bv[i] = b[i];
    sh.y[i] = 0.0;
  }
  __syncthreads();
  for (int i = tid; i < m; i += BT) {
    const int a = sh.cp[i], z = sh.cp[i + 1];
    int s = 0;
    for (int e = a; e < z; ++e) {
      sh.con[e] = (unsigned char)i;
      s += (sh.er[e] == sh.ec[e]) ? 1 : 2;
    }
    sh.nz[i] = ((double)(z - a) >= (double)n * n * 0.20) ? n * n : s;
  }
  __syncthreads();
  // sort keys (position, constraint, entry) are distinct: the ranks are a
  // permutation whatever the data (repeated entries just add up)
  int* key = reinterpret_cast<int*>(sh.S);
  for (int e = tid; e < ne; e += BT)
    key[e] = ((sh.er[e] + n * sh.ec[e]) * MMAX + sh.con[e]) * kEnt + e;
  __syncthreads();
  for (int e = tid; e < ne; e += BT) {
    const int ke = key[e];
    int rank = 0;
    for (int f = 0; f < ne; ++f) rank += key[f] < ke;
    sh.perm[rank] = (unsigned short)e;
  }
  __syncthreads();
  for (int s = tid; s < ne; s += BT) {
    const int e = sh.perm[s];
    int h = 1;
    if (s > 0) {
      const int f = sh.perm[s - 1];
      h = (sh.er[f] != sh.er[e] || sh.ec[f] != sh.ec[e]) ? 1 : 0;
    }
    sh.head[s] = (unsigned char)h;
  }
  __syncthreads();
  return ne;
}

// out_i = A_i . D (D global n x n; off-diagonal entries count twice)
template <int BT>
__device__ __forceinline__ void adot(int n, int m, const double* D, double* out, Shared<BT>& sh) {
  for (int i = threadIdx.x; i < m; i += BT) {
    double s = 0.0;
    for (int e = sh.cp[i]; e < sh.cp[i + 1]; ++e) {
      const int r = sh.er[e], c = sh.ec[e];
      double d = D[r + n * c];
      if (r != c) d += D[c + n * r];
      s += sh.ev[e] * d;
    }
    out[i] = s;
  }
  __syncthreads();
}

// out = sg sum_i yv_i A_i + c1 M1 + c2 M2 (global n x n; M1, M2 may be
// null).  Each position is summed by one thread in constraint order.
template <int BT>
__device__ __forceinline__ void yA_comb(int n, int ne, const double* yv, double sg, const double* M1, double c1,
                        const double* M2, double c2, double* out, Shared<BT>& sh) {
  const int nn = n * n;
  for (int e = threadIdx.x; e < nn; e += BT) {
    double v = 0.0;
    if (M1) v += c1 * M1[e];
    if (M2) v += c2 * M2[e];
    out[e] = v;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < ne; s += BT) {
    if (!sh.head[s]) continue;
    const int e = sh.perm[s];
    const int r = sh.er[e], c = sh.ec[e];
    double v = 0.0;
    for (int t = s; t < ne; ++t) {
      if (t > s && sh.head[t]) break;
      const int f = sh.perm[t];
      v += yv[sh.con[f]] * sh.ev[f];
    }
    out[r + n * c] += sg * v;
    if (r != c) out[c + n * r] += sg * v;
  }
  __syncthreads();
}

// acc = op(A) op(B) on the 32x32 quarter at (r0, c0) of a 64x64 product,
// A and B in LDS (ld LDA, zero beyond the data up to kd).  The MFMA's A
// operand is fed from B and its B operand from A, so the 16x16 result is the
// transpose: acc[ti][tj][v] is row r0 + 16 ti + lane%16, column
// c0 + 16 tj + 4 v + lane/16 (f64 16x16x4: D[i][j] in lane j + 16 (i%4),
// register i/4).
__device__ __forceinline__ void mma_quarter(v4d acc[2][2], const double* As, int ta,
                                            const double* Bs, int tb, int kd, int r0, int c0) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int kk = 0; kk < kd; kk += 4) {
    const int k = kk + lk;
    double a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int col = c0 + 16 * t + li, row = r0 + 16 * t + li;
      a[t] = tb ? Bs[col + LDA * k] : Bs[k + LDA * col];
      b[t] = ta ? As[k + LDA * row] : As[row + LDA * k];
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
        acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[tj], b[ti], acc[ti][tj], 0, 0, 0);
  }
}

// C = op(As) op(Bs) (LDS operands, n x n) into global C (ld n), or into sh.A
// (ld LDA) when C is null.  Operands must be zero in rows/columns [n, 64).
template <int BT>
__device__ __forceinline__ void mma_lds(int n, const double* As, int ta, const double* Bs, int tb, double* C,
                        Shared<BT>& sh) {
  constexpr int NW = BT / 64, NQ = 4 / NW;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int kd = (n + 3) & ~3;
  v4d acc[NQ][2][2];
#pragma unroll
  for (int s = 0; s < NQ; ++s) {
    const int q = w + NW * s, r0 = 32 * (q & 1), c0 = 32 * (q >> 1);
    if (r0 < n && c0 < n) {
      mma_quarter(acc[s], As, ta, Bs, tb, kd, r0, c0);
    } else {
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) acc[s][ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
    }
  }
  if (!C) __syncthreads();  // the operands may be sh.A
  double* out = C ? C : sh.A;
  const int ldo = C ? n : LDA;
#pragma unroll
  for (int s = 0; s < NQ; ++s) {
    const int q = w + NW * s, r0 = 32 * (q & 1), c0 = 32 * (q >> 1);
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + 16 * ti + (lane & 15), col = c0 + 16 * tj + 4 * v + (lane >> 4);
          if (row < n && col < n) out[row + ldo * col] = acc[s][ti][tj][v];
        }
  }
  __syncthreads();
}

// an n x n global matrix into LDS (ld LDA), zero outside it up to the extent
// the product's quarters read (32 or 64)
template <int BT>
__device__ __forceinline__ void load_pad(int n, const double* G, double* L) {
  const int R = n <= 32 ? 32 : 64;
  for (int e = threadIdx.x; e < R * R; e += BT) {
    const int i = e % R, j = e / R;
    L[i + LDA * j] = (i < n && j < n) ? G[i + n * j] : 0.0;
  }
}

// C = op(A) op(B), A, B global n x n; C global or (null) sh.A
template <int BT>
__device__ __forceinline__ void gemm(int n, const double* A, int ta, const double* B, int tb, double* C,
                     Shared<BT>& sh) {
  load_pad<BT>(n, A, sh.A);
  load_pad<BT>(n, B, sh.B2);
  __syncthreads();
  mma_lds(n, sh.A, ta, sh.B2, tb, C, sh);
}

// Cholesky of the Schur matrix in sh.S with rATL_dpotrfL's pivot rule
// (sdpa_dpotrf.cpp:218-265: a pivot below -1e-6 fails, one below 1e-14
// becomes 1e100).  Right-looking with one barrier per column: column k stays
// unscaled until the end and each update forms the scaled values it needs,
// S_ij -= (S_ik / L_kk)(S_jk / L_kk), the division as a multiplication by
// 1 / L_kk as the reference BLAS dtrsm does.  Four elements per thread per
// pass, their loads issued together.  On success S holds L (lower) and
// sh.piv[k] = 1 / L_kk.  Returns 0 or k + 1.
template <int BT>
__device__ __forceinline__ int chol_schur(int m, Shared<BT>& sh) {
  double* S = sh.S;
  for (int k = 0; k < m; ++k) {
    double d = uni(S[k + LDB * k]);
    if (d < -1.0e-6) return k + 1;
    if (d < 1.0e-14) d = 1.0e100;
    const double s = sqrt(d), rs = 1.0 / s;
    if (threadIdx.x == 0) sh.piv[k] = s;
    const int rem = m - k - 1, tot = rem * rem;
    for (int e0 = threadIdx.x; e0 < tot; e0 += 4 * BT) {
      double sij[4], lik[4], ljk[4];
      int idx[4];
      bool low[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = min(e0 + u * BT, tot - 1);
        const int i = k + 1 + e % rem, j = k + 1 + e / rem;
        idx[u] = i + LDB * j;
        low[u] = i >= j;
        sij[u] = S[idx[u]];
        lik[u] = S[i + LDB * k];
        ljk[u] = S[j + LDB * k];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (e0 + u * BT < tot && low[u]) S[idx[u]] = sij[u] - (lik[u] * rs) * (ljk[u] * rs);
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < m; k += BT) {
    const double s = sh.piv[k], rs = 1.0 / s;
    S[k + LDB * k] = s;
    for (int i = k + 1; i < m; ++i) S[i + LDB * k] *= rs;
    sh.piv[k] = rs;
  }
  __syncthreads();
  return 0;
}

// x := L^-T L^-1 x with the Schur factor (one wave; lane l keeps x_l and
// x_{l+64} in registers, the step's pivot value comes by readlane)
template <int BT>
__device__ __forceinline__ void schur_solve(int m, double* x, Shared<BT>& sh) {
  if (threadIdx.x < 64) {
    const int l = threadIdx.x;
    const double* S = sh.S;
    double x0 = l < m ? x[l] : 0.0, x1 = l + 64 < m ? x[l + 64] : 0.0;
    const double p0 = l < m ? sh.piv[l] : 0.0, p1 = l + 64 < m ? sh.piv[l + 64] : 0.0;
    // the column (row) of L each step needs is loaded unconditionally (rows
    // past m read the padded array and are masked by a select), so the loads
    // do not wait on the step's pivot value
    const int l1 = (l + 64 < MMAX) ? l + 64 : MMAX - 1;
    for (int k = 0; k < m; ++k) {
      const double t0 = S[l + LDB * k], t1 = S[l1 + LDB * k];
      const double xk = k < 64 ? rdlane(x0, k) * rdlane(p0, k) : rdlane(x1, k - 64) * rdlane(p1, k - 64);
      x0 = (l == k) ? xk : ((l > k && l < m) ? x0 - t0 * xk : x0);
      x1 = (l + 64 == k) ? xk : ((l + 64 > k && l + 64 < m) ? x1 - t1 * xk : x1);
    }
    for (int k = m - 1; k >= 0; --k) {
      const double t0 = S[k + LDB * l], t1 = S[k + LDB * l1];
      const double xk = k < 64 ? rdlane(x0, k) * rdlane(p0, k) : rdlane(x1, k - 64) * rdlane(p1, k - 64);
      x0 = (l == k) ? xk : ((l < k) ? x0 - t0 * xk : x0);
      x1 = (l + 64 == k) ? xk : ((l + 64 < k) ? x1 - t1 * xk : x1);
    }
    if (l < m) x[l] = x0;
    if (l + 64 < m) x[l + 64] = x1;
  }
  __syncthreads();
}

// Precision 1: the Schur matrix B stays in the lower triangle of sh.S (FP64,
// for the refinement residuals); the FP32 Cholesky factor of the
// equilibrated matrix s_i B_ij s_j (s_i = B_ii^-1/2, unit diagonal) takes the
// strictly upper slots -- L_ij (i > j) as a float in the first half of the
// double slot S[j + LDB i] -- and its pivots sh.fd.  Right-looking with one
// barrier per column, LAPACK's rule (a pivot that is not > 0 fails: the
// caller then factors B in FP64 with SDPA's rule).  Returns 0 or k + 1.
template <int BT>
__device__ __forceinline__ float& f32_slot(Shared<BT>& sh, int i, int j) {
  return *reinterpret_cast<float*>(&sh.S[j + LDB * i]);  // i > j
}

template <int BT>
__device__ __forceinline__ int chol_schur32(int m, Shared<BT>& sh) {
  double* S = sh.S;
  for (int i = threadIdx.x; i < m; i += BT) {
    const double d = S[i + LDB * i];
    sh.eqs[i] = d > 0.0 ? 1.0 / sqrt(d) : 1.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m * m; e += BT) {
    const int i = e % m, j = e / m;
    if (i > j) f32_slot(sh, i, j) = (float)(sh.eqs[i] * S[i + LDB * j] * sh.eqs[j]);
    else if (i == j) sh.fd[i] = (float)(sh.eqs[i] * S[i + LDB * i] * sh.eqs[i]);
  }
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    const float d = sh.fd[k];
    if (!(d > 0.0f)) return k + 1;
    const float rs = 1.0f / sqrtf(d);
    const int rem = m - k - 1, tot = rem * rem;
    // the lower triangle (i >= j > k) of the trailing matrix
    for (int e = threadIdx.x; e < tot; e += BT) {
      const int ii = k + 1 + e % rem, jj = k + 1 + e / rem;
      if (ii < jj) continue;
      const float lik = f32_slot(sh, ii, k) * rs, ljk = f32_slot(sh, jj, k) * rs;
      if (ii == jj) sh.fd[ii] -= lik * ljk;
      else f32_slot(sh, ii, jj) -= lik * ljk;
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < m; k += BT) {
    const float s = sqrtf(sh.fd[k]), rs = 1.0f / s;
    for (int i = k + 1; i < m; ++i) f32_slot(sh, i, k) *= rs;
    sh.fd[k] = s;
  }
  __syncthreads();
  return 0;
}

// x := (L L^T)^-1 x with the FP32 factor (one wave, FP32 arithmetic; lane l
// keeps x_l and x_{l+64})
template <int BT>
__device__ __forceinline__ void schur_solve32(int m, double* x, Shared<BT>& sh) {
  if (threadIdx.x < 64) {
    const int l = threadIdx.x;
    float x0 = l < m ? (float)x[l] : 0.0f, x1 = l + 64 < m ? (float)x[l + 64] : 0.0f;
    const float p0 = l < m ? 1.0f / sh.fd[l] : 0.0f, p1 = l + 64 < m ? 1.0f / sh.fd[l + 64] : 0.0f;
    for (int k = 0; k < m; ++k) {
      const float t0 = l > k && l < m ? f32_slot(sh, l, k) : 0.0f;
      const float t1 = l + 64 > k && l + 64 < m ? f32_slot(sh, l + 64, k) : 0.0f;
      const float xk = k < 64 ? __shfl(x0, k) * __shfl(p0, k) : __shfl(x1, k - 64) * __shfl(p1, k - 64);
      x0 = (l == k) ? xk : x0 - t0 * xk;
      x1 = (l + 64 == k) ? xk : x1 - t1 * xk;
    }
    for (int k = m - 1; k >= 0; --k) {
      const float t0 = l < k ? f32_slot(sh, k, l) : 0.0f;
      const float t1 = l + 64 < k ? f32_slot(sh, k, l + 64) : 0.0f;
      const float xk = k < 64 ? __shfl(x0, k) * __shfl(p0, k) : __shfl(x1, k - 64) * __shfl(p1, k - 64);
      x0 = (l == k) ? xk : x0 - t0 * xk;
      x1 = (l + 64 == k) ? xk : x1 - t1 * xk;
    }
    if (l < m) x[l] = x0;
    if (l + 64 < m) x[l + 64] = x1;
  }
  __syncthreads();
}

// x := B^-1 x by the FP32 factor of the equilibrated matrix and FP64
// iterative refinement: r = x0 - B x in FP64 (B's lower triangle), until
// LAPACK dsposv's test on the equilibrated system holds,
// |S r|_inf <= sqrt(m) eps |S B S|_inf |S^-1 x|_inf.  A residual that does
// not halve is a stall: returns 1 (x unchanged), else 0.  *steps: the
// refinement steps taken.
template <int BT>
__device__ __forceinline__ int schur_refine(int m, double* x, Shared<BT>& sh, int* steps) {
  double* S = sh.S;
  // |S B S|_inf (every thread gets it)
  double bn = 0.0;
  for (int i = threadIdx.x; i < m; i += BT) {
    double a = 0.0;
    for (int j = 0; j < m; ++j) a += fabs(S[max(i, j) + LDB * min(i, j)]) * sh.eqs[j];
    bn = fmax(bn, a * sh.eqs[i]);
  }
  {
    double o = 0.0;
    bmax2(bn, o, sh);
  }
  for (int i = threadIdx.x; i < m; i += BT) {
    sh.rr[i] = x[i];
    sh.xr[i] = 0.0;
  }
  __syncthreads();
  double prev = -1.0;
  for (int it = 0; it < 30; ++it) {
    for (int i = threadIdx.x; i < m; i += BT) sh.rr[i] *= sh.eqs[i];
    __syncthreads();
    schur_solve32(m, sh.rr, sh);
    for (int i = threadIdx.x; i < m; i += BT) sh.xr[i] += sh.eqs[i] * sh.rr[i];
    __syncthreads();
    // r = x0 - B xr, and the norms of the test
    double rn = 0.0, xn = 0.0;
    for (int i = threadIdx.x; i < m; i += BT) {
      double a = x[i];
      for (int j = 0; j < m; ++j) a -= S[max(i, j) + LDB * min(i, j)] * sh.xr[j];
      sh.rr[i] = a;
      rn = fmax(rn, fabs(a) * sh.eqs[i]);
      xn = fmax(xn, fabs(sh.xr[i]) / sh.eqs[i]);
    }
    bmax2(rn, xn, sh);
    ++*steps;
    if (rn <= sqrt((double)m) * 1.1102230246251565e-16 * bn * xn) {
      for (int i = threadIdx.x; i < m; i += BT) x[i] = sh.xr[i];
      __syncthreads();
      return 0;
    }
    if (prev >= 0.0 && rn > 0.5 * prev) return 1;
    prev = rn;
  }
  return 1;
}

// M (LDS, ld LDA, symmetric n x n, lower part read) := L^-1 (lower) where
// M = L L^T, with LAPACK's rule (a pivot that is not > 0 fails), by a group
// of NS waves (ws = this wave's index in the group).  Every wave of the
// work-group calls at once: the groups for X and Z pass the same barriers.
// Lane l owns row l; wave ws updates the columns j = ws (mod NS).  Column k
// stays unscaled during the factorization (the waves read it while others
// update theirs) and each update forms L_lk L_jk from it, the column values
// of the step coming by readlane; one barrier per column.  The inverse is
// dtrti2's column sweep from the right on wave 0 of the group: column j of
// L^-1 is -(L^-1)_{22} L_{21} / L_jj with the trailing inverse in place.
// M must have kPadCols writable columns behind its 64.  pv: n doubles of
// LDS for the pivots.  Returns 0 or k + 1.
template <int NS>
__device__ __forceinline__ int chol_inv_group(int n, double* M, double* pv, int ws) {
  const int l = threadIdx.x & 63;
  int info = 0;
  for (int k = 0; k < n; ++k) {
    const double d = uni(M[k + LDA * k]);
    if (!info && !(d > 0.0)) info = k + 1;
    const double s = info ? 1.0 : sqrt(d), r = 1.0 / s;
    const double li = (l > k && l < n) ? M[l + LDA * k] * r : 0.0;
    if (ws == 0 && l == k) pv[k] = s;
    // owned columns past k, eight at a time, straight-line: the loads are
    // in flight together, rows l < j take a zero update, and columns past
    // 63 fall in the pad columns behind the matrix (lane l >= n holds
    // li = 0, so columns in [n, 64) are rewritten unchanged)
    const int jf = k + 1 + (((ws - (k + 1)) % NS) + NS) % NS;
    for (int j0 = jf; j0 < n; j0 += 8 * NS) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = M[l + LDA * (j0 + NS * u)];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = j0 + NS * u;
        const double upd = li * rdlane(li, j & 63);
        M[l + LDA * j] = t[u] - ((l >= j) ? upd : 0.0);
      }
    }
    if (NS > 1) __syncthreads();
  }
  // scale the columns: L_lk = M_lk / L_kk (as the multiplication above)
  for (int k = ws; k < n; k += NS) {
    const double s = pv[k], r = 1.0 / s;
    if (l > k && l < n) M[l + LDA * k] *= r;
    if (l == k) M[k + LDA * k] = s;
  }
  if (NS > 1) __syncthreads();
  if (ws == 0 && !info) {
    for (int j = n - 1; j >= 0; --j) {
      const double ajj = 1.0 / M[j + LDA * j];
      const double x = (l > j && l < n) ? M[l + LDA * j] : 0.0;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      for (int k0 = j + 1; k0 < n; k0 += 8) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = M[l + LDA * (k0 + u)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int k = k0 + u;
          const double xk = rdlane(x, k & 63);
          acc[u & 3] += (l >= k && k < n) ? t[u] * xk : 0.0;
        }
      }
      if (l > j && l < n) M[l + LDA * j] = -ajj * ((acc[0] + acc[1]) + (acc[2] + acc[3]));
      if (l == j) M[j + LDA * j] = ajj;
    }
  }
  return info;
}

// calF3 with A_j the owner (outer loop), sdpa_newton.cpp:1319-1403, over
// the owner's entries e1 = first, first + step, ...; X and Z^-1 in LDS
template <int BT>
__device__ __forceinline__ double calF3(const Shared<BT>& sh, int i, int j, int first, int step) {
  const double* X = sh.A;
  const double* Zi = sh.B2;
  double ret = 0.0;
  for (int e1 = sh.cp[j] + first; e1 < sh.cp[j + 1]; e1 += step) {
    const int a = sh.er[e1], b = sh.ec[e1];
    const double* Xa = X + LDA * a;
    const double* Xb = X + LDA * b;
    const double* Za = Zi + LDA * a;
    const double* Zb = Zi + LDA * b;
    double sum = 0.0;
    for (int e2 = sh.cp[i]; e2 < sh.cp[i + 1]; ++e2) {
      const int g = sh.er[e2], d = sh.ec[e2];
      double plu;
      if (a != b) {
        plu = (g != d) ? Zb[d] * Xa[g] + Zb[g] * Xa[d] + Za[d] * Xb[g] + Za[g] * Xb[d]
                       : Zb[d] * Xa[g] + Za[d] * Xb[g];
      } else {
        plu = (g != d) ? Zb[d] * Xa[g] + Zb[g] * Xa[d] : Zb[d] * Xa[g];
      }
      sum += sh.ev[e2] * plu;
    }
    ret += sh.ev[e1] * sum;
  }
  return ret;
}


// the owner of pair (i, j): the larger NonZeroEffect (ties: the smaller index)
template <int BT>
__device__ __forceinline__ bool j_owns(const Shared<BT>& sh, int i, int j) {
  const int jnz = sh.nz[j], inz = sh.nz[i];
  return !((jnz < inz) || (inz == jnz && i < j));
}

// the Schur matrix into sh.S (both triangles), F3 with SDPA's ownership
// rule: light pairs one thread each, heavy pairs one wave each (the lanes
// split the owner's entries)
template <int BT>
__device__ __forceinline__ void make_bmat(int n, int m, const double* X, const double* Zi,
                                          int kHeavy, Shared<BT>& sh) {
  for (int e = threadIdx.x; e < n * n; e += BT) {
    const int i = e % n, j = e / n;
    sh.A[i + LDA * j] = X[e];
    sh.B2[i + LDA * j] = Zi[e];
  }
  __syncthreads();
  const int mm = m * m;
  for (int t = threadIdx.x; t < mm; t += BT) {
    const int i = t % m, j = t / m;
    if (i < j) continue;
    if ((sh.cp[i + 1] - sh.cp[i]) * (sh.cp[j + 1] - sh.cp[j]) > kHeavy) continue;
    const double v = j_owns(sh, i, j) ? calF3(sh, i, j, 0, 1) : calF3(sh, j, i, 0, 1);
    sh.S[i + LDB * j] = v;
    sh.S[j + LDB * i] = v;
  }
  const int lane = threadIdx.x & 63;
  for (int t0 = (threadIdx.x >> 6) * 64; t0 < mm; t0 += BT) {
    const int t = t0 + lane;
    bool heavy = false;
    if (t < mm) {
      const int i = t % m, j = t / m;
      heavy = i >= j && (sh.cp[i + 1] - sh.cp[i]) * (sh.cp[j + 1] - sh.cp[j]) > kHeavy;
    }
    unsigned long long mask = __ballot(heavy);
    while (mask) {
      const int tt = t0 + __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      const int i = tt % m, j = tt / m;
      const double part = j_owns(sh, i, j) ? calF3(sh, i, j, lane, 64) : calF3(sh, j, i, lane, 64);
      const double v = wsum(part);
      if (lane == 0) {
        sh.S[i + LDB * j] = v;
        sh.S[j + LDB * i] = v;
      }
    }
  }
  __syncthreads();
}

// smallest eigenvalue of the symmetric n x n (n <= 20) matrix in sh.A, as
// the dsyev branch of Jal::getMinEigen (sdpa_jordan.cpp:80-87, lower
// triangle): Householder reduction to tridiagonal form on wave 0 (lane l
// owns row l), then the tridiagonal's smallest eigenvalue by Sturm
// multisection on the whole work-group.
template <int BT>
__device__ __forceinline__ double small_min_eig(int n, Shared<BT>& sh) {
  if (threadIdx.x < 64) householder_tridiag_wave(n, sh.A, LDA, sh.dg, sh.od);
  __syncthreads();
  double th = 0.0, zl = 0.0;
  tridiag_min_block(n, sh.dg, sh.od, th, zl, BT, sh.trisc, false);
  __syncthreads();
  return th;
}

// Lal::getMinEigen (sdpa_linear.cpp:29-142) on the symmetric n x n (n > 20)
// matrix in sh.A: the truncated Lanczos recurrence with its stopping rules.
template <int BT>
__device__ __forceinline__ double lanczos(int n, Shared<BT>& sh) {
  const int tid = threadIdx.x;
  const double* M = sh.A;
  double mn = 1.0e+51, mn_old = 1.0e+52, error = 1.0e+10;
  int k = 0, kk = 0;
  for (int i = tid; i < n; i += BT) {
    sh.lq[i] = 0.0;
    sh.lr[i] = 1.0;
  }
  double beta = sqrt((double)n);
  const double sqn = sqrt((double)n);
  while (k < n && k < sqn + 10 && k < kLz && beta > 1.0e-16 &&
         (fabs(mn - mn_old) > (1.0e-5) * fabs(mn) + (1.0e-8) ||
          fabs(error * beta) > (1.0e-2) * fabs(mn) + (1.0e-4))) {
    const double inv = 1.0 / beta;
    __syncthreads();
    for (int i = tid; i < n; i += BT) {
      sh.lqo[i] = sh.lq[i];
      sh.lq[i] = sh.lr[i] * inv;
    }
    __syncthreads();
    for (int i = tid; i < n; i += BT) {
      double a0 = 0.0, a1 = 0.0;
      int j = 0;
#pragma unroll 4
      for (; j + 1 < n; j += 2) {
        a0 += M[i + LDA * j] * sh.lq[j];
        a1 += M[i + LDA * (j + 1)] * sh.lq[j + 1];
      }
      if (j < n) a0 += M[i + LDA * j] * sh.lq[j];
      sh.lw[i] = a0 + a1;
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
      if (tid == 0) {
        sh.sc[0] = th;
        sh.sc[1] = zl;
      }
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

// minimum eigenvalue of L D L^T (StepLength::MehrotraCorrector's
// Jal::getMinEigen(invChol, D)); T is global scratch
template <int BT>
__device__ __forceinline__ double min_eig(int n, const double* L, const double* D, double* T, Shared<BT>& sh) {
  gemm(n, D, 0, L, 1, T, sh);        // T = D L^T
  gemm(n, L, 0, T, 0, nullptr, sh);  // sh.A = L T
  return n <= 20 ? small_min_eig(n, sh) : lanczos(n, sh);
}

// invChol of X and Z (LXi, LZi, both lower) and Zi = LZi^T LZi: the two
// factorizations run at once on the two halves of the work-group.  0, or 1
// if one failed.
template <int BT>
__device__ __forceinline__ int inv_chol(int n, const double* X, const double* Z, double* LXi, double* LZi,
                        double* Zi, Shared<BT>& sh) {
  const int tid = threadIdx.x, w = tid >> 6;
  for (int e = tid; e < n * n; e += BT) {
    const int i = e % n, j = e / n;
    sh.A[i + LDA * j] = X[e];
    sh.B2[i + LDA * j] = Z[e];
  }
  __syncthreads();
  {
    // waves [0, NS) factor X, waves [NS, 2 NS) factor Z
    constexpr int NS = BT / 128;
    const int grp = w / NS, ws = w % NS;
    const int r = chol_inv_group<NS>(n, grp ? sh.B2 : sh.A, sh.pv[grp], ws);
    if ((tid & 63) == 0 && ws == 0) sh.info[grp] = r;
  }
  __syncthreads();
  if (sh.info[0] || sh.info[1]) return 1;
  for (int e = tid; e < NMAX * NMAX; e += BT) {
    const int i = e & 63, j = e >> 6;
    const bool in = i < n && j < n;
    const double vx = (in && i >= j) ? sh.A[i + LDA * j] : 0.0;
    const double vz = (in && i >= j) ? sh.B2[i + LDA * j] : 0.0;
    if (in) {
      LXi[i + n * j] = vx;
      LZi[i + n * j] = vz;
    }
    sh.B2[i + LDA * j] = vz;
  }
  __syncthreads();
  mma_lds(n, sh.B2, 1, sh.B2, 0, Zi, sh);  // Zi = LZi^T LZi
  return 0;
}

template <int BT>
__global__ __launch_bounds__(BT) void k_batch_solve(BData Dt, BParams Pm) {
  __shared__ Shared<BT> sh;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) sh.task = (int)atomicAdd(Dt.next, 1u);
    __syncthreads();
    const int slot = __builtin_amdgcn_readfirstlane(sh.task);
    __syncthreads();
    if (slot >= Dt.count) break;
    const int k = Dt.order[slot];
    const int n = Dt.n[k], m = Dt.m[k], nn = n * n;
    const double* Cm = Dt.C + Dt.coff[k];
    const int ne = stage(Dt, k, n, m, sh);
    double* wk = Dt.woff ? Dt.work + Dt.woff[k]
                         : Dt.work + (size_t)blockIdx.x * kMats * NMAX * NMAX;
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
    __syncthreads();
    const double muInit = lam * lam;
    double muCur = muInit;
    // residuals: p = b - A(X), D = C - sum y A - Z, and their max norms
    auto residuals = [&](double& nP, double& nD) {
      adot(n, m, X, sh.p, sh);
      for (int i = tid; i < m; i += BT) sh.p[i] = sh.bv[i] - sh.p[i];
      yA_comb(n, ne, sh.y, -1.0, Cm, 1.0, Z, -1.0, Dm, sh);
      double a = 0.0, c = 0.0;
      for (int i = tid; i < m; i += BT) a = fmax(a, fabs(sh.p[i]));
      for (int e = tid; e < nn; e += BT) c = fmax(c, fabs(Dm[e]));
      bmax2(a, c, sh);
      nP = a;
      nD = c;
    };
    double normP, normD;
    residuals(normP, normD);
    const double initNormP = normP, initNormD = normD;
    double betaVal = Pm.betaStar;
    double thetaP = (initNormP < Pm.epsilonDash) ? 0.0 : 1.0;
    double thetaD = (initNormD < Pm.epsilonDash) ? 0.0 : 1.0;
    double rho = 1.0;
    double etaP = Pm.omegaStar * nDim * muInit, etaD = etaP;
    double objP, objD;
    {
      double o[2] = {0.0, 0.0};
      for (int e = tid; e < nn; e += BT) o[0] += Cm[e] * X[e];
      for (int i = tid; i < m; i += BT) o[1] += sh.bv[i] * sh.y[i];
      bsums<BT, 2>(o, sh);
      objP = o[0];
      objD = o[1];
    }
    int phase = noINFO, failure = 0;
    double alphaP = 0.0, alphaD = 0.0;
    // precision 1: FP32 factor + FP64 refinement until mu < muSwitch or the
    // first refinement stall (then FP64 for the rest of the instance's run)
    bool fp64_only = Pm.precision != 1;
    int nf32 = 0, nf64 = 0, nref = 0, nstall = 0;
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
        if (gap / dom <= Pm.epsilonStar) {
          phase = pdOPT;
          return false;
        }
      }
      if (v == noINFO && rho > 1.0 + NONZERO) {
        phase = pdINF;
        return false;
      }
      if (v == pFEAS) {
        if (objP <= -Pm.upperBound) {
          phase = pUNBD;
          return false;
        }
        if (rho > 1.0 + NONZERO) {
          phase = pFEAS_dINF;
          return false;
        }
      }
      if (v == dFEAS) {
        if (objD >= -Pm.lowerBound) {
          phase = dUNBD;
          return false;
        }
        if (rho > 1.0 + NONZERO) {
          phase = pINF_dFEAS;
          return false;
        }
      }
      return true;
    };
    update_phase();
    // r_zinv = target Zinv - X (- Dx Dz Zinv), minus X D Zinv when
    // infeasible; g = -A(.) + p   (Newton::Make_gVec)
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
      adot(n, m, Dd, sh.g, sh);
      for (int i = tid; i < m; i += BT) sh.g[i] = -sh.g[i] + sh.p[i];
      __syncthreads();
    };
    // Dz = -sum dy A (+ D); Dx = sym(r_zinv - X Dz Zinv)
    auto make_direction = [&](int infeasible) {
      yA_comb(n, ne, sh.dy, -1.0, infeasible ? Dm : nullptr, 1.0, nullptr, 0.0, Dz, sh);
      gemm(n, X, 0, Dz, 0, T, sh);
      gemm(n, T, 0, Zi, 0, W, sh);
      for (int e = tid; e < nn; e += BT) {
        const int i = e % n, j = e / n;
        if (i > j) {
          const double s = ((R[i + n * j] - W[i + n * j]) + (R[j + n * i] - W[j + n * i])) * 0.5;
          Dx[i + n * j] = s;
          Dx[j + n * i] = s;
        } else if (i == j) {
          Dx[e] = R[e] - W[e];
        }
      }
      __syncthreads();
    };
    // q = C.Dx, b.dy, X.Dz, Dx.Z, Dx.Dz, b.y, C.X (the last two when all)
    auto direction_dots = [&](double (&q)[7], bool all) {
      double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int e = tid; e < nn; e += BT) {
        const double dx = Dx[e], dz = Dz[e];
        v[0] += Cm[e] * dx;
        v[2] += X[e] * dz;
        v[3] += dx * Z[e];
        v[4] += dx * dz;
        if (all) v[6] += Cm[e] * X[e];
      }
      for (int i = tid; i < m; i += BT) {
        v[1] += sh.bv[i] * sh.dy[i];
        if (all) v[5] += sh.bv[i] * sh.y[i];
      }
      bsums<BT, 7>(v, sh);
      for (int t = 0; t < 7; ++t) q[t] = v[t];
    };
    int it = 0;
    while (update_phase() && it < Pm.maxIteration) {
      const bool affine = !(phase == noINFO || phase == pFEAS || phase == dFEAS);
      if (phase == pdFEAS) betaVal = 0.0;
      else betaVal = affine ? 2.0 : Pm.betaBar;
      const int infeasible = (phase == pFEAS || phase == noINFO) ? 1 : 0;
      // Newton::Mehrotra(PREDICTOR)
      make_gvec(0, betaVal * muCur, infeasible);
      make_bmat(n, m, X, Zi, Pm.heavy, sh);
      bool use32 = !fp64_only && muCur >= Pm.muSwitch;
      if (use32 && chol_schur32(m, sh)) use32 = false;  // B is intact: factor it in FP64
      if (use32) {
        ++nf32;
      } else {
        ++nf64;
        if (chol_schur(m, sh)) {
          failure = 1;
          break;
        }
      }
      // dy := B^-1 g; a refinement stall falls back to the FP64 factorization
      auto schur_dy = [&]() -> bool {
        for (int i = tid; i < m; i += BT) sh.dy[i] = sh.g[i];
        __syncthreads();
        if (use32) {
          if (!schur_refine(m, sh.dy, sh, &nref)) return true;
          ++nstall;
          ++nf64;
          use32 = false;
          fp64_only = true;
          if (chol_schur(m, sh)) return false;
        }
        schur_solve(m, sh.dy, sh);
        return true;
      };
      if (!schur_dy()) {
        failure = 1;
        break;
      }
      make_direction(infeasible);
      double q[7];
      direction_dots(q, false);
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
      if (!schur_dy()) {
        failure = 1;
        break;
      }
      make_direction(infeasible);
      // StepLength::MehrotraCorrector
      const double lamX = min_eig(n, LXi, Dx, T, sh);
      const double lamZ = min_eig(n, LZi, Dz, T, sh);
      const double alphaBD = 100.0;
      alphaP = (-lamX > 1.0 / alphaBD) ? -1.0 / lamX : alphaBD;
      alphaD = (-lamZ > 1.0 / alphaBD) ? -1.0 / lamZ : alphaBD;
      alphaP *= Pm.gammaStar;
      alphaD *= Pm.gammaStar;
      direction_dots(q, true);
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
        if (maxRatio > 0.0 && maxRatio < 1.0) {
          alphaP *= maxRatio;
          alphaD *= maxRatio;
        }
      }
      // Solutions::update (sdpa_dataset.cpp:145-171)
      for (int e = tid; e < nn; e += BT) {
        X[e] += alphaP * Dx[e];
        Z[e] += alphaD * Dz[e];
      }
      for (int i = tid; i < m; i += BT) sh.y[i] += alphaD * sh.dy[i];
      __syncthreads();
      if (alphaP < 1.0e-4 && alphaD < 1.0e-4) {
        failure = 2;
        break;
      }
      if (inv_chol(n, X, Z, LXi, LZi, Zi, sh)) {
        failure = 3;
        break;
      }
      if (!affine) {
        thetaP = fabs((1.0 - alphaP) * thetaP);
        thetaD = fabs((1.0 - alphaD) * thetaD);
      }
      double o[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // trZ, trX, X.Z, C.X, b.y
      for (int i = tid; i < n; i += BT) {
        o[0] += Z[i + n * i];
        o[1] += X[i + n * i];
      }
      for (int e = tid; e < nn; e += BT) {
        o[2] += X[e] * Z[e];
        o[3] += Cm[e] * X[e];
      }
      for (int i = tid; i < m; i += BT) o[4] += sh.bv[i] * sh.y[i];
      bsums<BT, 5>(o, sh);
      objP = o[3];
      objD = o[4];
      muCur = o[2] / nDim;
      residuals(normP, normD);
      thetaP = (initNormP > Pm.epsilonDash * 1.0e-2) ? normP / initNormP : 0.0;
      thetaD = (initNormD > Pm.epsilonDash * 1.0e-2) ? normD / initNormD : 0.0;
      // SolveInfo::update (sdpa_parts.cpp:914-1060)
      rho = 0.0;
      const double x0z0 = nDim * muInit, xzn = nDim * muCur;
      const double x0z = lam * o[0], xz0 = lam * o[1];
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
      if (Dt.stats) {
        double* st = Dt.stats + 4 * (size_t)k;
        st[0] = nf32;
        st[1] = nf64;
        st[2] = nref;
        st[3] = nstall;
      }
    }
    __syncthreads();
  }
}

}  // namespace
}  // namespace sdpa_amd

using namespace sdpa_amd;

// Checks of the instance arrays shared by both entry points (the error codes
// of include/sdpa_batch.h); fills the largest-first queue order and, when
// woff is given, the per-instance slab offsets (returns their total).
static int check_batch(int count, const int* n, const int* m, const long* coff, long nC,
                       const long* boff, long nb, const long* pptr, const int* cptr, long ncptr,
                       const int* er, const int* ec, long nent, const double* lambda,
                       std::vector<int>& order, std::vector<long>* woff, long* wtot,
                       int input_order = 0) {
  long tot = 0;
  if (woff) woff->assign(count, 0);
  for (int k = 0; k < count; ++k) {
    if (n[k] < 1 || n[k] > NMAX || m[k] < 1 || m[k] > MMAX) return -3;
    if (coff[k] < 0 || coff[k] + (long)n[k] * n[k] > nC || boff[k] < 0 || boff[k] + m[k] > nb ||
        pptr[k] < 0 || pptr[k] + m[k] + 1 > ncptr)
      return -4;
    for (int i = 0; i <= m[k]; ++i) {
      const int e = cptr[pptr[k] + i];
      if (e < 0 || e > nent || (i > 0 && e < cptr[pptr[k] + i - 1])) return -5;
    }
    if (cptr[pptr[k] + m[k]] - cptr[pptr[k]] > kEnt) return -7;
    if (!(lambda[k] > 0.0)) return -8;
    if (woff) (*woff)[k] = tot;
    tot += (long)kMats * n[k] * n[k];
  }
  for (int k = 0; k < count; ++k)
    for (int e = cptr[pptr[k]]; e < cptr[pptr[k] + m[k]]; ++e)
      if (er[e] < 0 || ec[e] < er[e] || ec[e] >= n[k]) return -6;
  // largest instances first (work ~ m^3 / 3 + n^3 per iteration): the
  // queue then ends on small ones and the work-groups finish together
  order.resize(count);
  std::iota(order.begin(), order.end(), 0);
  if (input_order) {
    if (wtot) *wtot = tot;
    return 0;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    const double wa = (double)m[a] * m[a] * m[a] / 3.0 + (double)n[a] * n[a] * n[a];
    const double wc = (double)m[c] * m[c] * m[c] / 3.0 + (double)n[c] * n[c] * n[c];
    return wa > wc;
  });
  if (wtot) *wtot = tot;
  return 0;
}

extern "C" int sdpa_batch_solve_cfg(int device, int count, const int* n, const int* m,
                                    const long* coff, const double* C, long nC, const long* boff,
                                    const double* b, long nb, const long* pptr, const int* cptr,
                                    long ncptr, const int* er, const int* ec, const double* ev,
                                    long nent, const double* lambda, const double* param,
                                    const sdpa_batch_config* cfg, double* out, double* stats,
                                    double* kernel_ms) {
  if (count <= 0 || !n || !m || !coff || !C || !boff || !b || !pptr || !cptr || !er || !ec ||
      !ev || !lambda || !param || !out || !cfg)
    return -1;
  const int threads = cfg->threads;
  if (threads != 128 && threads != 256) return -2;
  if (cfg->heavy < 1 || (cfg->order != 0 && cfg->order != 1) ||
      (cfg->precision != 0 && cfg->precision != 1) || !(cfg->mu_switch >= 0.0))
    return -10;
  std::vector<long> woff;
  std::vector<int> order;
  long wtot = 0;
  if (int e = check_batch(count, n, m, coff, nC, boff, nb, pptr, cptr, ncptr, er, ec, nent,
                          lambda, order, &woff, &wtot, cfg->order))
    return e;
  if (hipSetDevice(device) != hipSuccess) return -100;
  hipStream_t s;
  if (hipStreamCreate(&s) != hipSuccess) return -100;
  int *dn = nullptr, *dm = nullptr, *dcp = nullptr, *der = nullptr, *dec = nullptr;
  int* dord = nullptr;
  unsigned* dnext = nullptr;
  long *dco = nullptr, *dbo = nullptr, *dpp = nullptr, *dwo = nullptr;
  double *dC = nullptr, *db = nullptr, *dev = nullptr, *dl = nullptr, *dw = nullptr, *dout = nullptr;
  double* dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
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
    if ((rc = up((void**)&dord, order.data(), sizeof(int) * count))) break;
    if (hipMalloc(&dnext, sizeof(unsigned)) != hipSuccess ||
        hipMemsetAsync(dnext, 0, sizeof(unsigned), s) != hipSuccess) {
      rc = -100;
      break;
    }
    if (hipMalloc(&dw, sizeof(double) * (wtot ? wtot : 1)) != hipSuccess) {
      rc = -100;
      break;
    }
    if (hipMalloc(&dout, sizeof(double) * 8 * count) != hipSuccess) {
      rc = -100;
      break;
    }
    if (stats && hipMalloc(&dst, sizeof(double) * 4 * count) != hipSuccess) {
      rc = -100;
      break;
    }
    BData D{count, dn, dm, dco, dC, dbo, db, dpp, dcp, der, dec, dev, dl, dwo, dord, dnext, dw,
            dout, dst};
    BParams P{(int)param[0], param[1], param[2], param[3], param[4], param[5], param[6], param[7],
              param[8], cfg->heavy, cfg->precision, cfg->mu_switch};
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
      rc = -100;
      break;
    }
    // persistent work-groups, one per CU (the LDS use allows one), taking
    // instances from the queue until it is empty
    const int grid = count < ncu ? count : ncu;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      rc = -100;
      break;
    }
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
    if (hipMemcpy(out, dout, sizeof(double) * 8 * count, hipMemcpyDeviceToHost) != hipSuccess)
      rc = -100;
    if (stats && hipMemcpy(stats, dst, sizeof(double) * 4 * count, hipMemcpyDeviceToHost) !=
                     hipSuccess)
      rc = -100;
  } while (0);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* p : {(void*)dn, (void*)dm, (void*)dcp, (void*)der, (void*)dec, (void*)dco, (void*)dbo,
                  (void*)dpp, (void*)dwo, (void*)dC, (void*)db, (void*)dev, (void*)dl, (void*)dw,
                  (void*)dout, (void*)dord, (void*)dnext, (void*)dst})
    if (p) (void)hipFree(p);
  (void)hipStreamDestroy(s);
  return rc;
}

extern "C" int sdpa_batch_solve(int device, int count, const int* n, const int* m,
                                const long* coff, const double* C, long nC, const long* boff,
                                const double* b, long nb, const long* pptr, const int* cptr,
                                long ncptr, const int* er, const int* ec, const double* ev,
                                long nent, const double* lambda, const double* param,
                                int threads, double* out, double* kernel_ms) {
  const sdpa_batch_config cfg{threads, kDefaultHeavy, 0, 0, 0.0};
  return sdpa_batch_solve_cfg(device, count, n, m, coff, C, nC, boff, b, nb, pptr, cptr, ncptr,
                              er, ec, ev, nent, lambda, param, &cfg, out, nullptr, kernel_ms);
}

// Launch-configuration search: every candidate of the grid (work-group
// size x heavy-pair threshold x queue order) runs the whole solve of a
// sample of the batch (every stride-th instance, at least 64 and at most
// `sample` of them) twice, the faster kernel time of the two counting;
// the fastest candidate is returned.  The precision and mu switch are the
// caller's and are not searched (they change the results, the launch shape
// does not change what converges).
extern "C" int sdpa_batch_autotune(int device, int count, const int* n, const int* m,
                                   const long* coff, const double* C, long nC, const long* boff,
                                   const double* b, long nb, const long* pptr, const int* cptr,
                                   long ncptr, const int* er, const int* ec, const double* ev,
                                   long nent, const double* lambda, const double* param,
                                   int precision, double mu_switch, int sample,
                                   sdpa_batch_config* best, double* table, int table_cap) {
  if (count <= 0 || !best || sample < 1) return -1;
  // the sample: instances k = stride * i, renumbered with their own offsets
  const int ns = std::min(count, std::max(64, sample));
  const int stride = std::max(1, count / ns);
  std::vector<int> sn, sm, scp, ser, sec;
  std::vector<long> sco, sbo, spp;
  std::vector<double> sC, sb, sev, slam;
  for (int i = 0; i < ns && (long)i * stride < count; ++i) {
    const int k = i * stride;
    if (n[k] < 1 || n[k] > NMAX || m[k] < 1 || m[k] > MMAX || coff[k] < 0 ||
        coff[k] + (long)n[k] * n[k] > nC || boff[k] < 0 || boff[k] + m[k] > nb || pptr[k] < 0 ||
        pptr[k] + m[k] + 1 > ncptr)
      return -4;
    sn.push_back(n[k]);
    sm.push_back(m[k]);
    sco.push_back((long)sC.size());
    sC.insert(sC.end(), C + coff[k], C + coff[k] + (long)n[k] * n[k]);
    sbo.push_back((long)sb.size());
    sb.insert(sb.end(), b + boff[k], b + boff[k] + m[k]);
    spp.push_back((long)scp.size());
    const int e0 = cptr[pptr[k]], e1 = cptr[pptr[k] + m[k]];
    if (e0 < 0 || e1 < e0 || e1 > nent) return -5;
    for (int q = 0; q <= m[k]; ++q) scp.push_back(cptr[pptr[k] + q] - e0 + (int)sev.size());
    ser.insert(ser.end(), er + e0, er + e1);
    sec.insert(sec.end(), ec + e0, ec + e1);
    sev.insert(sev.end(), ev + e0, ev + e1);
    slam.push_back(lambda[k]);
  }
  const int cnt = (int)sn.size();
  std::vector<double> out(8 * (size_t)cnt);
  const int thr[2] = {256, 128};
  const int hv[5] = {16, 32, 64, 128, 256};
  double bestms = -1.0;
  int row = 0;
  for (int a = 0; a < 2; ++a)
    for (int h = 0; h < 5; ++h)
      for (int o = 0; o < 2; ++o) {
        const sdpa_batch_config cfg{thr[a], hv[h], o, precision, mu_switch};
        double ms = -1.0;
        for (int rep = 0; rep < 2; ++rep) {
          double t = 0.0;
          const int rc = sdpa_batch_solve_cfg(
              device, cnt, sn.data(), sm.data(), sco.data(), sC.data(), (long)sC.size(),
              sbo.data(), sb.data(), (long)sb.size(), spp.data(), scp.data(), (long)scp.size(),
              ser.data(), sec.data(), sev.data(), (long)sev.size(), slam.data(), param, &cfg,
              out.data(), nullptr, &t);
          if (rc) return rc;
          ms = (ms < 0.0 || t < ms) ? t : ms;
        }
        if (table && row < table_cap) {
          double* r = table + 4 * (size_t)row;
          r[0] = cfg.threads;
          r[1] = cfg.heavy;
          r[2] = cfg.order;
          r[3] = ms;
        }
        ++row;
        if (bestms < 0.0 || ms < bestms) {
          bestms = ms;
          *best = cfg;
        }
      }
  return row;
}

// ---------------------------------------------------------------------------
// The same solve with no device allocation: every device buffer of the
// launch (the staged instance data, the queue, the per-work-group slabs and
// the results) lives in one fixed device array, batch_scratch.hip's (a code
// object of its own, so that only callers of this entry point load it).
namespace sdpa_amd {
size_t noalloc_scratch_bytes();
int noalloc_scratch_address(int device, void** out);
}  // namespace sdpa_amd

namespace {

__global__ void k_zero_u32(unsigned* p) {
  if (threadIdx.x == 0) *p = 0u;
}

// bump allocator over the scratch array (256-byte aligned pieces)
struct Carve {
  unsigned char* base;
  size_t used = 0;
  template <class T>
  T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base + used);
    used += (sizeof(T) * (count ? count : 1) + 255) & ~(size_t)255;
    return p;
  }
};

}  // namespace

extern "C" long sdpa_batch_noalloc_capacity(void) { return (long)noalloc_scratch_bytes(); }

extern "C" int sdpa_batch_solve_noalloc(int device, int count, const int* n, const int* m,
                                        const long* coff, const double* C, long nC,
                                        const long* boff, const double* b, long nb,
                                        const long* pptr, const int* cptr, long ncptr,
                                        const int* er, const int* ec, const double* ev,
                                        long nent, const double* lambda, const double* param,
                                        int threads, double* out, void* stream) {
  if (count <= 0 || !n || !m || !coff || !C || !boff || !b || !pptr || !cptr || !er || !ec ||
      !ev || !lambda || !param || !out)
    return -1;
  if (threads != 128 && threads != 256) return -2;
  if (device < 0) return -100;
  std::vector<int> order;
  if (int e = check_batch(count, n, m, coff, nC, boff, nb, pptr, cptr, ncptr, er, ec, nent,
                          lambda, order, nullptr, nullptr))
    return e;
  if (hipSetDevice(device) != hipSuccess) return -100;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return -100;
  const int grid = count < ncu ? count : ncu;
  // the layout of this call in the scratch array; reject what does not fit
  Carve cv{nullptr};
  const size_t o_n = cv.used; cv.take<int>(count);
  const size_t o_m = cv.used; cv.take<int>(count);
  const size_t o_co = cv.used; cv.take<long>(count);
  const size_t o_C = cv.used; cv.take<double>(nC);
  const size_t o_bo = cv.used; cv.take<long>(count);
  const size_t o_b = cv.used; cv.take<double>(nb);
  const size_t o_pp = cv.used; cv.take<long>(count);
  const size_t o_cp = cv.used; cv.take<int>(ncptr);
  const size_t o_er = cv.used; cv.take<int>(nent);
  const size_t o_ec = cv.used; cv.take<int>(nent);
  const size_t o_ev = cv.used; cv.take<double>(nent);
  const size_t o_l = cv.used; cv.take<double>(count);
  const size_t o_ord = cv.used; cv.take<int>(count);
  const size_t o_next = cv.used; cv.take<unsigned>(1);
  const size_t o_out = cv.used; cv.take<double>(8 * (size_t)count);
  const size_t o_w = cv.used; cv.take<double>((size_t)grid * kMats * NMAX * NMAX);
  if (cv.used > noalloc_scratch_bytes()) return -9;
  void* base = nullptr;
  if (noalloc_scratch_address(device, &base)) return -100;
  unsigned char* B0 = static_cast<unsigned char*>(base);
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto put = [&](size_t off, const void* src, size_t bytes) -> int {
    if (bytes && hipMemcpyAsync(B0 + off, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
      return -100;
    return 0;
  };
  int rc = 0;
  if ((rc = put(o_n, n, sizeof(int) * count)) || (rc = put(o_m, m, sizeof(int) * count)) ||
      (rc = put(o_co, coff, sizeof(long) * count)) || (rc = put(o_C, C, sizeof(double) * nC)) ||
      (rc = put(o_bo, boff, sizeof(long) * count)) || (rc = put(o_b, b, sizeof(double) * nb)) ||
      (rc = put(o_pp, pptr, sizeof(long) * count)) || (rc = put(o_cp, cptr, sizeof(int) * ncptr)) ||
      (rc = put(o_er, er, sizeof(int) * nent)) || (rc = put(o_ec, ec, sizeof(int) * nent)) ||
      (rc = put(o_ev, ev, sizeof(double) * nent)) ||
      (rc = put(o_l, lambda, sizeof(double) * count)) ||
      (rc = put(o_ord, order.data(), sizeof(int) * count)))
    return rc;
  // the queue head starts at zero on every call (a kernel, not a memset, so
  // that the call can sit in a captured graph)
  unsigned* dnext = reinterpret_cast<unsigned*>(B0 + o_next);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, dnext);
  BData D{count,
          reinterpret_cast<const int*>(B0 + o_n),
          reinterpret_cast<const int*>(B0 + o_m),
          reinterpret_cast<const long*>(B0 + o_co),
          reinterpret_cast<const double*>(B0 + o_C),
          reinterpret_cast<const long*>(B0 + o_bo),
          reinterpret_cast<const double*>(B0 + o_b),
          reinterpret_cast<const long*>(B0 + o_pp),
          reinterpret_cast<const int*>(B0 + o_cp),
          reinterpret_cast<const int*>(B0 + o_er),
          reinterpret_cast<const int*>(B0 + o_ec),
          reinterpret_cast<const double*>(B0 + o_ev),
          reinterpret_cast<const double*>(B0 + o_l),
          nullptr,  // per-work-group slabs
          reinterpret_cast<const int*>(B0 + o_ord),
          dnext,
          reinterpret_cast<double*>(B0 + o_w),
          reinterpret_cast<double*>(B0 + o_out),
          nullptr};
  BParams P{(int)param[0], param[1], param[2], param[3], param[4], param[5], param[6], param[7],
            param[8], kDefaultHeavy, 0, 0.0};
  if (threads == 256)
    hipLaunchKernelGGL(k_batch_solve<256>, dim3(grid), dim3(256), 0, s, D, P);
  else
    hipLaunchKernelGGL(k_batch_solve<128>, dim3(grid), dim3(128), 0, s, D, P);
  if (hipGetLastError() != hipSuccess) return -101;
  if (hipMemcpyAsync(out, B0 + o_out, sizeof(double) * 8 * count, hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return -100;
  if (hipStreamSynchronize(s) != hipSuccess) return -101;
  return 0;
}
