// eig_dev.h -- device routines shared by the step-length kernels (eig.hip)
// and the batched small-SDP solver (batch.hip): wave sums, an FP64
// reciprocal, the smallest eigenvalue of a symmetric tridiagonal matrix with
// the last component of its eigenvector (what the reference reads from
// dsteqr, sdpa_linear.cpp:98-127), and the Householder reduction to
// tridiagonal form for small blocks (the dsyev path, sdpa_linear.cpp:145-182).
#pragma once
#include <hip/hip_runtime.h>

namespace sdpa_amd {

// lane l's value of v (l the same in every lane): two v_readlane, no LDS
// round trip as a general shuffle takes
__device__ __forceinline__ double rdlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v from another lane of the same row of 16 by a DPP pattern (both halves)
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Reductions over a full wave (all 64 lanes active), the same value in every
// lane: DPP within rows of 16 (quad swaps [1,0,3,2] and [2,3,0,1], half-row
// mirror, row mirror), then the four row results by readlane -- no LDS
// round trips, as a shuffle butterfly takes.
__device__ __forceinline__ double wsum(double v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return (rdlane(v, 0) + rdlane(v, 16)) + (rdlane(v, 32) + rdlane(v, 48));
}

__device__ __forceinline__ double wmin(double v) {
  v = fmin(v, dpp_mov<0xB1>(v));
  v = fmin(v, dpp_mov<0x4E>(v));
  v = fmin(v, dpp_mov<0x141>(v));
  v = fmin(v, dpp_mov<0x140>(v));
  return fmin(fmin(rdlane(v, 0), rdlane(v, 16)), fmin(rdlane(v, 32), rdlane(v, 48)));
}

__device__ __forceinline__ double wmax(double v) {
  v = fmax(v, dpp_mov<0xB1>(v));
  v = fmax(v, dpp_mov<0x4E>(v));
  v = fmax(v, dpp_mov<0x141>(v));
  v = fmax(v, dpp_mov<0x140>(v));
  return fmax(fmax(rdlane(v, 0), rdlane(v, 16)), fmax(rdlane(v, 32), rdlane(v, 48)));
}

// 1/x to full FP64 accuracy without the IEEE divide sequence: the hardware
// reciprocal estimate and two Newton steps (each doubles the correct bits).
// x is a nonzero normal number here (Sturm pivots are kept away from 0).
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Smallest eigenvalue theta of the symmetric tridiagonal T (diagonal d[0..N),
// off-diagonal e[0..N-1), in LDS) and the last component of its unit
// eigenvector -- what the reference reads from dsteqr's output (the first
// eigenvalue in ascending order and Q[k][0], sdpa_linear.cpp:98-127).
// All threads of the work-group (nthreads, a multiple of 64) take part:
// multisection with Sturm counts, nthreads shifts per round (the interval of
// the smallest eigenvalue shrinks nthreads+1-fold per round); then wave 0 runs
// two steps of inverse iteration from a shift just below theta, where
// T - sigma I is positive definite and its LDL^T needs no pivoting.
// The multisection stops at a relative width of 2 eps, or at an absolute
// width of 1e-3 eps ||T|| (far below the eps ||T|| error of a backward-stable
// QR step, and reached in about 8 rounds when theta is near zero).
// scratch: 2 * (nthreads / 64) doubles of LDS.  Call from all threads.
// want_z = false skips the inverse iteration (zlast is then 0).
static __device__ __forceinline__ void tridiag_min_block(int N, const double* d, const double* e,
                                                         double& theta, double& zlast,
                                                         int nthreads, double* scratch,
                                                         bool want_z = true) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = nthreads >> 6;
  const double eps = 2.220446049250313e-16;
  double lo = 1.0e300, hi = -1.0e300;
  for (int i = lane; i < N; i += 64) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < N - 1 ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
  }
  lo = wmin(lo);
  hi = wmax(hi);
  const double tnorm = fmax(fabs(lo), fabs(hi));
  const double pivmin = 1.0e-290 + eps * eps * tnorm * tnorm;
  lo -= 4.0 * eps * tnorm + pivmin;
  hi += 4.0 * eps * tnorm + pivmin;
  int* firsts = reinterpret_cast<int*>(scratch);
  for (int round = 0; round < 20; ++round) {
    const double step = (hi - lo) / (double)(nthreads + 1);
    const double x = lo + step * (double)(tid + 1);
    // Sturm count: eigenvalues below x
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
#pragma unroll 4
    for (int i = 1; i < N; ++i) {
      q = (d[i] - x) - (e[i - 1] * e[i - 1]) * recip(q);
      if (fabs(q) < pivmin) q = -pivmin;
      cnt += q < 0.0;
    }
    // the first shift (over the whole work-group) with at least one
    // eigenvalue below it brackets the smallest eigenvalue
    const unsigned long long mask = __ballot(cnt >= 1);
    if (lane == 0) firsts[wv] = mask ? (__ffsll((long long)mask) - 1) + 64 * wv : nthreads;
    __syncthreads();
    int first = nthreads;
    for (int w = 0; w < nw; ++w) first = min(first, firsts[w]);
    __syncthreads();
    const double nlo = lo + step * (double)first;
    const double nhi = (first < nthreads) ? lo + step * (double)(first + 1) : hi;
    lo = nlo;
    hi = nhi;
    if (hi - lo <= fmax(2.0 * eps * fmax(fabs(lo), fabs(hi)), 1.0e-3 * eps * tnorm) + pivmin) break;
  }
  theta = 0.5 * (lo + hi);
  zlast = 0.0;
  if (!want_z) return;
  // inverse iteration on wave 0: LDL^T of T - sigma I (pivots piv_i), two
  // solves from b = ones with a normalization between them.  The sequential
  // sweeps run on every lane with the same scalars; lane i keeps entry i.
  if (wv == 0) {
    const double sigma = theta - (8.0 * eps * tnorm + pivmin);
    double piv = 0.0, rpiv = 0.0, y = 0.0, zi = 0.0;
    double rprev = 1.0, yprev = 0.0;
    for (int i = 0; i < N; ++i) {  // pivots and forward sweep L y = 1
      const double ei = (i > 0) ? e[i - 1] : 0.0;
      double p = (d[i] - sigma) - ((i > 0) ? ei * ei * rprev : 0.0);
      if (fabs(p) < pivmin) p = pivmin;
      const double rp = recip(p);
      const double l = (i > 0) ? ei * rprev : 0.0;
      const double yi = 1.0 - l * yprev;
      if (lane == i) { piv = p; rpiv = rp; y = yi; }
      rprev = rp;
      yprev = yi;
    }
    (void)piv;
    // backward: z_i = (y_i - e_i z_{i+1}) / piv_i
    double znext = 0.0;
    for (int i = N - 1; i >= 0; --i) {
      const double ri = rdlane(rpiv, i), yi = rdlane(y, i);
      const double ei = (i < N - 1) ? e[i] : 0.0;
      const double zi2 = (yi - ei * znext) * ri;
      if (lane == i) zi = zi2;
      znext = zi2;
    }
    // normalize and second solve with b = z
    double s2 = (lane < N) ? zi * zi : 0.0;
    s2 = wsum(s2);
    const double b = zi / sqrt(s2);
    rprev = 1.0;
    yprev = 0.0;
    for (int i = 0; i < N; ++i) {
      const double ei = (i > 0) ? e[i - 1] : 0.0;
      const double l = (i > 0) ? ei * rprev : 0.0;
      const double yi = rdlane(b, i) - l * yprev;
      if (lane == i) y = yi;
      rprev = rdlane(rpiv, i);
      yprev = yi;
    }
    znext = 0.0;
    for (int i = N - 1; i >= 0; --i) {
      const double ri = rdlane(rpiv, i), yi = rdlane(y, i);
      const double ei = (i < N - 1) ? e[i] : 0.0;
      const double zi2 = (yi - ei * znext) * ri;
      if (lane == i) zi = zi2;
      znext = zi2;
    }
    double nrm2 = (lane < N) ? zi * zi : 0.0;
    nrm2 = wsum(nrm2);
    const double last = rdlane(zi, N - 1);
    zlast = last / sqrt(nrm2);
  }
}

// Householder reduction of the symmetric n x n matrix A (LDS, leading
// dimension lda; the lower triangle is read and mirrored first) to
// tridiagonal form: diagonal dg[0..n), off-diagonal od[0..n-1).  One wave,
// lane l holds row l (n <= 64); A is overwritten.  Columns already
// tridiagonal are skipped.
static __device__ __forceinline__ void householder_tridiag_wave(int n, double* A, int lda,
                                                                double* dg, double* od) {
  const int l = threadIdx.x & 63;
  for (int j = 0; j < n; ++j)
    if (l < j) A[l + lda * j] = A[j + lda * l];
  for (int k = 0; k + 2 < n; ++k) {
    const bool act = l > k && l < n;
    const double xl = act ? A[l + lda * k] : 0.0;
    const double x0 = rdlane(xl, k + 1);
    const double tail = wsum((l > k + 1 && l < n) ? xl * xl : 0.0);
    if (tail == 0.0) {  // already tridiagonal in this column
      if (l == 0) {
        dg[k] = A[k + lda * k];
        od[k] = x0;
      }
      continue;
    }
    const double alpha = -copysign(sqrt(x0 * x0 + tail), x0);
    const double vl = (l == k + 1) ? x0 - alpha : xl;
    const double tau = 2.0 / (tail + (x0 - alpha) * (x0 - alpha));
    double pl = 0.0;
    for (int j = k + 1; j < n; ++j) {
      const double vj = rdlane(vl, j);
      if (act) pl += A[l + lda * j] * vj;
    }
    pl *= tau;
    const double K = 0.5 * tau * wsum(pl * vl);
    const double wl = pl - K * vl;
    for (int j = k + 1; j < n; ++j) {
      const double vj = rdlane(vl, j), wj = rdlane(wl, j);
      if (act) A[l + lda * j] -= vl * wj + wl * vj;
    }
    if (l == 0) {
      dg[k] = A[k + lda * k];
      od[k] = alpha;
    }
  }
  if (l == 0) {
    if (n >= 2) {
      dg[n - 2] = A[(n - 2) + lda * (n - 2)];
      od[n - 2] = A[(n - 1) + lda * (n - 2)];
    }
    dg[n - 1] = A[(n - 1) + lda * (n - 1)];
  }
}


}  // namespace sdpa_amd
