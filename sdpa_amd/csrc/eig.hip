// eig.hip -- step-length eigenvalue estimates on the device
// (StepLength::computeStepLength -> Jal::getMinEigen, sdpa_parts.cpp:363-432,
// sdpa_jordan.cpp:58-109).
//
// For each SDP block the scaled direction M = L D L^T (L = inverse Cholesky
// factor, D = Dx or Dz) is formed by two batched MFMA GEMMs.  One wave per
// (block, X/Z) then runs
//   n > 20 : the truncated Lanczos of Lal::getMinEigen (sdpa_linear.cpp:29-143):
//            start vector of ones, at most min(n, sqrt(n)+10) steps, the
//            tridiagonal eigenproblem re-solved on the schedule of line 84,
//            the stop tests of lines 52-56, result min - |err*beta|;
//   n <= 20: all eigenvalues of M by cyclic Jacobi (the dsyev path, 145-182).
// The tridiagonal eigenproblem is solved by implicit QL with Wilkinson shifts
// (Bowdler-Martin-Reinsch-Wilkinson 1968), eigenvectors kept one row per lane.
#include "ctx.h"

#include <cmath>

namespace sdpa_amd {

constexpr int kMaxLanczos = 64;   // tridiagonal size limit (sqrt(n)+10 <= 64 for n <= 2916)
constexpr int kMaxN = 2048;       // largest block the single-wave kernel holds in LDS vectors

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
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
// scratch: 2 * (nthreads / 64) doubles of LDS.  Call from all threads.
__device__ void tridiag_min_block(int N, const double* d, const double* e, double& theta,
                                  double& zlast, int nthreads, double* scratch) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = nthreads >> 6;
  const double eps = 2.220446049250313e-16;
  double lo = 1.0e300, hi = -1.0e300;
  for (int i = lane; i < N; i += 64) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < N - 1 ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
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
    if (hi - lo <= 2.0 * eps * fmax(fabs(lo), fabs(hi)) + pivmin) break;
  }
  theta = 0.5 * (lo + hi);
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
      const double ri = __shfl(rpiv, i, 64), yi = __shfl(y, i, 64);
      const double ei = (i < N - 1) ? e[i] : 0.0;
      const double zi2 = (yi - ei * znext) * ri;
      if (lane == i) zi = zi2;
      znext = zi2;
    }
    // normalize and second solve with b = z
    double s2 = (lane < N) ? zi * zi : 0.0;
    for (int o = 32; o > 0; o >>= 1) s2 += __shfl_xor(s2, o, 64);
    const double b = zi / sqrt(s2);
    rprev = 1.0;
    yprev = 0.0;
    for (int i = 0; i < N; ++i) {
      const double ei = (i > 0) ? e[i - 1] : 0.0;
      const double l = (i > 0) ? ei * rprev : 0.0;
      const double yi = __shfl(b, i, 64) - l * yprev;
      if (lane == i) y = yi;
      rprev = __shfl(rpiv, i, 64);
      yprev = yi;
    }
    znext = 0.0;
    for (int i = N - 1; i >= 0; --i) {
      const double ri = __shfl(rpiv, i, 64), yi = __shfl(y, i, 64);
      const double ei = (i < N - 1) ? e[i] : 0.0;
      const double zi2 = (yi - ei * znext) * ri;
      if (lane == i) zi = zi2;
      znext = zi2;
    }
    double nrm2 = (lane < N) ? zi * zi : 0.0;
    for (int o = 32; o > 0; o >>= 1) nrm2 += __shfl_xor(nrm2, o, 64);
    const double last = __shfl(zi, N - 1, 64);
    zlast = last / sqrt(nrm2);
  }
}

// Smallest eigenvalue of a symmetric n x n (n <= 64) matrix in LDS (ld 65),
// cyclic Jacobi; lane k updates row/column k.
__device__ double jacobi_min_wave(int n, double* A) {
  const int lane = threadIdx.x;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int j = 0; j < n; ++j) {
      double v = (lane < n && lane > j) ? A[lane + 65 * j] : 0.0;
      off += v * v;
    }
    off = wsum(off);
    double dv = lane < n ? A[lane + 65 * lane] : 0.0;
    dg = wsum(dv * dv);
    // converged when the off-diagonal norm is below eps times the diagonal's
    // (a tighter test is never met in FP64 and ran every call to the sweep cap)
    if (off == 0.0 || off <= 1e-31 * dg) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        __syncthreads();
        double apq = A[p + 65 * q];
        if (apq == 0.0) continue;
        double app = A[p + 65 * p], aqq = A[q + 65 * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        __syncthreads();
        if (lane < n) {
          double akp = A[lane + 65 * p], akq = A[lane + 65 * q];
          A[lane + 65 * p] = c * akp - s * akq;
          A[lane + 65 * q] = s * akp + c * akq;
        }
        __syncthreads();
        if (lane < n) {
          double apk = A[p + 65 * lane], aqk = A[q + 65 * lane];
          A[p + 65 * lane] = c * apk - s * aqk;
          A[q + 65 * lane] = s * apk + c * aqk;
        }
      }
  }
  __syncthreads();
  double dv = lane < n ? A[lane + 65 * lane] : 1.0e300;
  for (int o = 32; o > 0; o >>= 1) dv = fmin(dv, __shfl_xor(dv, o, 64));
  return dv;
}

// n <= 20: one wave per (block, which), cyclic Jacobi on M = L D L^T.
// Ms: packed scaled directions for X (which 0) and Z (which 1), at offsets 0
// and `stride`.
__global__ __launch_bounds__(64) void k_min_eig_small(const double* __restrict__ Ms, long stride,
                                                      const long* __restrict__ off,
                                                      const int* __restrict__ nn, int nblk,
                                                      double* __restrict__ out) {
  const int q = blockIdx.x, which = blockIdx.y;
  const int n = nn[q];
  if (n > 20) return;
  const double* M = Ms + which * stride + off[q];
  const int lane = threadIdx.x;
  __shared__ double Zs[64 * 65];
  for (int j = 0; j < n; ++j)
    if (lane < n) Zs[lane + 65 * j] = M[lane + (size_t)n * j];
  // dsyev('L') reads the lower triangle: mirror it
  __syncthreads();
  for (int j = 0; j < n; ++j)
    if (lane < n && lane < j) Zs[lane + 65 * j] = Zs[j + 65 * lane];
  __syncthreads();
  const double result = jacobi_min_wave(n, Zs);
  if (lane == 0) out[which * nblk + q] = result;
}

constexpr int kLzThreads = 256;

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wsum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < kLzThreads / 64; ++i) s += red[i];
  return s;
}

// n > 20: the truncated Lanczos of Lal::getMinEigen on M, 256 threads per
// (block, which); in the matrix-vector product each thread owns rows, the
// tridiagonal eigenproblem runs on wave 0.
__global__ __launch_bounds__(kLzThreads) void k_lanczos(const double* __restrict__ Ms, long stride,
                                                        const long* __restrict__ off,
                                                        const int* __restrict__ nn, int nblk,
                                                        double* __restrict__ out) {
  const int q = blockIdx.x, which = blockIdx.y;
  const int n = nn[q];
  if (n <= 20) return;
  const double* M = Ms + which * stride + off[q];
  const int tid = threadIdx.x;
  __shared__ double dg[kMaxLanczos], od[kMaxLanczos];  // tridiagonal of the recurrence
  __shared__ double qv[kMaxN], qold[kMaxN], rv[kMaxN], wv[kMaxN];
  __shared__ double red[kLzThreads / 64];
  __shared__ double sh[4];  // theta, last eigenvector component
  __shared__ double trisc[2 * (kLzThreads / 64)];
  double mn = 1.0e+51, mn_old = 1.0e+52, error = 1.0e+10;
  int k = 0, kk = 0;
  for (int i = tid; i < n; i += kLzThreads) { qv[i] = 0.0; rv[i] = 1.0; }
  double beta = sqrt((double)n);
  const double sqn = sqrt((double)n);
  while (k < n && k < sqn + 10 && k < kMaxLanczos && beta > 1.0e-16 &&
         (fabs(mn - mn_old) > (1.0e-5) * fabs(mn) + (1.0e-8) ||
          fabs(error * beta) > (1.0e-2) * fabs(mn) + (1.0e-4))) {
    const double inv = 1.0 / beta;
    __syncthreads();
    for (int i = tid; i < n; i += kLzThreads) { qold[i] = qv[i]; qv[i] = rv[i] * inv; }
    __syncthreads();
    // w = M q: one thread per row i, lanes over consecutive rows, so each
    // step j reads 64 consecutive entries of column j (coalesced); q_j is an
    // LDS broadcast.  32 columns of loads in flight per thread, no cross-lane
    // reduction.
    for (int i = tid; i < n; i += kLzThreads) {
      const double* Mi = M + i;
      double a[4] = {0.0, 0.0, 0.0, 0.0};
      int j = 0;
      for (; j + 32 <= n; j += 32) {
        double mv[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) mv[u] = Mi[(size_t)n * (j + u)];
#pragma unroll
        for (int u = 0; u < 32; ++u) a[u & 3] += mv[u] * qv[j + u];
      }
      for (; j < n; ++j) a[j & 3] += Mi[(size_t)n * j] * qv[j];
      wv[i] = (a[0] + a[1]) + (a[2] + a[3]);
    }
    __syncthreads();
    double a = 0.0;
    for (int i = tid; i < n; i += kLzThreads) a += qv[i] * wv[i];
    const double alpha = block_sum(a, red);
    if (tid == 0) dg[k] = alpha;
    for (int i = tid; i < n; i += kLzThreads) rv[i] = (wv[i] - qv[i] * alpha) - qold[i] * beta;
    if (kk >= sqrt((double)k) || k == n - 1 || k > sqrt((double)n + 9)) {
      kk = 0;
      __syncthreads();
      {
        double th = 0.0, zl = 0.0;
        tridiag_min_block(k + 1, dg, od, th, zl, kLzThreads, trisc);
        if (tid == 0) { sh[0] = th; sh[1] = zl; }
      }
      __syncthreads();
      mn_old = mn;
      mn = sh[0];
      error = sh[1];
    }
    double v2 = 0.0;
    for (int i = tid; i < n; i += kLzThreads) v2 += rv[i] * rv[i];
    beta = sqrt(block_sum(v2, red));
    if (tid == 0) od[k] = beta;
    ++k;
    ++kk;
  }
  if (tid == 0) out[which * nblk + q] = mn - fabs(error * beta);
}

int lp_min_step(Ctx* c, const double* D, const double* L, double* out);

int min_eig(Ctx* c, double* lamX, double* lamZ) {
  const int nb = c->nSDP;
  double mX = 1.0e50, mZ = 1.0e50;
  if (nb > 0) {
    for (int q = 0; q < nb; ++q)
      if (c->sdpN[q] > kMaxN) return -7;
    // M = L D L^T for X (into w1) and Z (into w2): T = D L^T (w3), M = L T
    std::vector<GemmItem> it(4 * nb);
    int maxn = 0;
    for (int q = 0; q < nb; ++q) {
      const int n = c->sdpN[q];
      const long o = c->sdpOff[q];
      it[q] = GemmItem{c->d_Dx + o, c->d_invCholX + o, c->d_w3 + o, n, n, n, n, n, n};
      it[nb + q] = GemmItem{c->d_Dz + o, c->d_invCholZ + o, c->d_w3 + c->lpOff + o, n, n, n, n, n, n};
      it[2 * nb + q] = GemmItem{c->d_invCholX + o, c->d_w3 + o, c->d_w1 + o, n, n, n, n, n, n};
      it[3 * nb + q] = GemmItem{c->d_invCholZ + o, c->d_w3 + c->lpOff + o, c->d_w1 + c->lpOff + o,
                                n, n, n, n, n, n};
      maxn = n > maxn ? n : maxn;
    }
    if (nb == 1 && maxn >= kBlasMinN) {
      // one large block: plain GEMMs through the library
      for (int k = 0; k < 4; ++k) {
        const GemmItem& g = it[k];
        if (int e = blas_gemm(c, 0, k < 2 ? 1 : 0, g.M, g.N, g.K, 1.0, g.A, g.lda, g.B, g.ldb,
                              0.0, g.C, g.ldc))
          return e;
      }
    } else {
      if (int e = c->dw.ensure_items(4 * nb)) return e;
      SDPA_HIP_CHECK(hipMemcpyAsync(c->dw.d_items, it.data(), sizeof(GemmItem) * 4 * nb,
                                    hipMemcpyHostToDevice, c->stream));
      if (int e = gemm_batched(c->stream, c->dw.d_items, 2 * nb, maxn, maxn, 0, 1, 1.0, 0.0, 0))
        return e;
      if (int e = gemm_batched(c->stream, c->dw.d_items + 2 * nb, 2 * nb, maxn, maxn, 0, 0, 1.0,
                               0.0, 0))
        return e;
    }
    hipLaunchKernelGGL(k_min_eig_small, dim3(nb, 2), dim3(64), 0, c->stream, c->d_w1, c->lpOff,
                       c->d_sdpOff, c->d_sdpN, nb, c->d_eig);
    hipLaunchKernelGGL(k_lanczos, dim3(nb, 2), dim3(kLzThreads), 0, c->stream, c->d_w1, c->lpOff,
                       c->d_sdpOff, c->d_sdpN, nb, c->d_eig);
    SDPA_HIP_CHECK(hipGetLastError());
    std::vector<double> h(2 * nb);
    SDPA_HIP_CHECK(hipMemcpyAsync(h.data(), c->d_eig, sizeof(double) * 2 * nb,
                                  hipMemcpyDeviceToHost, c->stream));
    SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < nb; ++q) {
      mX = h[q] < mX ? h[q] : mX;
      mZ = h[nb + q] < mZ ? h[nb + q] : mZ;
    }
  }
  if (c->nLP > 0) {
    double a, b;
    if (int e = lp_min_step(c, c->d_Dx, c->d_invCholX, &a)) return e;
    if (int e = lp_min_step(c, c->d_Dz, c->d_invCholZ, &b)) return e;
    mX = a < mX ? a : mX;
    mZ = b < mZ ? b : mZ;
  }
  *lamX = mX;
  *lamZ = mZ;
  return 0;
}

}  // namespace sdpa_amd
