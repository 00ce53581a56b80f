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

// Implicit QL on a (N <= 64) tridiagonal held one entry per lane (dreg = d[lane],
// ereg = e[lane]); Z is N x N in LDS (ld 65), row = lane.  Eigenvalues end
// ascending in dreg.  Returns nonzero if an eigenvalue did not converge.
__device__ int ql_wave(int N, double& dreg, double& ereg, double* Z) {
  const int lane = threadIdx.x;
  for (int i = 0; i < N; ++i) Z[lane + 65 * i] = (lane == i) ? 1.0 : 0.0;
  if (lane == N - 1) ereg = 0.0;
  int fail = 0;
  for (int l = 0; l < N && !fail; ++l) {
    int iter = 0;
    for (;;) {
      int mm;
      for (mm = l; mm < N - 1; ++mm) {
        double dm = __shfl(dreg, mm, 64), dm1 = __shfl(dreg, mm + 1, 64);
        double em = __shfl(ereg, mm, 64);
        if (fabs(em) <= 2.220446049250313e-16 * (fabs(dm) + fabs(dm1))) break;
      }
      if (mm == l) break;
      if (++iter > 30 * N) { fail = 1; break; }
      double dl = __shfl(dreg, l, 64), dl1 = __shfl(dreg, l + 1, 64), el = __shfl(ereg, l, 64);
      double g = (dl1 - dl) / (2.0 * el);
      double r = hypot(g, 1.0);
      g = __shfl(dreg, mm, 64) - dl + el / (g + (g >= 0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      bool brk = false;
      for (i = mm - 1; i >= l; --i) {
        double ei = __shfl(ereg, i, 64);
        double f = s * ei, b = c * ei;
        r = hypot(f, g);
        if (lane == i + 1) ereg = r;
        if (r == 0.0) {
          if (lane == i + 1) dreg -= p;
          if (lane == mm) ereg = 0.0;
          brk = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = __shfl(dreg, i + 1, 64) - p;
        r = (__shfl(dreg, i, 64) - g) * s + 2.0 * c * b;
        p = s * r;
        if (lane == i + 1) dreg = g + p;
        g = c * r - b;
        if (lane < N) {
          double zf = Z[lane + 65 * (i + 1)];
          double zi = Z[lane + 65 * i];
          Z[lane + 65 * (i + 1)] = s * zi + c * zf;
          Z[lane + 65 * i] = c * zi - s * zf;
        }
      }
      if (brk && i >= l) continue;
      if (lane == l) { dreg -= p; ereg = g; }
      if (lane == mm) ereg = 0.0;
    }
  }
  // selection sort, ascending, with the eigenvector columns
  for (int i = 0; i < N - 1; ++i) {
    int kmin = i;
    double dmin = __shfl(dreg, i, 64);
    for (int j = i + 1; j < N; ++j) {
      double dj = __shfl(dreg, j, 64);
      if (dj < dmin) { dmin = dj; kmin = j; }
    }
    if (kmin != i) {
      double di = __shfl(dreg, i, 64);
      if (lane == i) dreg = dmin;
      if (lane == kmin) dreg = di;
      if (lane < N) {
        double z = Z[lane + 65 * i];
        Z[lane + 65 * i] = Z[lane + 65 * kmin];
        Z[lane + 65 * kmin] = z;
      }
    }
  }
  return fail;
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
    if (off == 0.0 || off <= 1e-36 * dg) break;
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

// One wave per (block, which).  Ms: packed scaled directions for X (which 0)
// and Z (which 1), stacked at offsets 0 and lpOff.
__global__ __launch_bounds__(64) void k_min_eig(const double* __restrict__ Ms, long stride,
                                                const long* __restrict__ off,
                                                const int* __restrict__ nn, int nblk,
                                                double* __restrict__ out) {
  const int q = blockIdx.x, which = blockIdx.y;
  const int n = nn[q];
  const double* M = Ms + which * stride + off[q];
  const int lane = threadIdx.x;
  __shared__ double Zs[64 * 65];
  __shared__ double qv[kMaxN], qold[kMaxN], rv[kMaxN], wv[kMaxN];
  double result;
  if (n <= 20) {
    for (int j = 0; j < n; ++j)
      if (lane < n) Zs[lane + 65 * j] = M[lane + (size_t)n * j];
    // dsyev('L') reads the lower triangle: mirror it
    __syncthreads();
    for (int j = 0; j < n; ++j)
      if (lane < n && lane < j) Zs[lane + 65 * j] = Zs[j + 65 * lane];
    __syncthreads();
    result = jacobi_min_wave(n, Zs);
  } else {
    double dreg = 1.0e50, ereg = 0.0;    // diag[lane], off[lane] of the Lanczos recurrence
    double mn = 1.0e+51, mn_old = 1.0e+52, error = 1.0e+10;
    int k = 0, kk = 0;
    for (int i = lane; i < n; i += 64) { qv[i] = 0.0; rv[i] = 1.0; }
    double beta = sqrt((double)n);
    const double sqn = sqrt((double)n);
    while (k < n && k < sqn + 10 && k < kMaxLanczos && beta > 1.0e-16 &&
           (fabs(mn - mn_old) > (1.0e-5) * fabs(mn) + (1.0e-8) ||
            fabs(error * beta) > (1.0e-2) * fabs(mn) + (1.0e-4))) {
      const double inv = 1.0 / beta;
      __syncthreads();
      for (int i = lane; i < n; i += 64) { qold[i] = qv[i]; qv[i] = rv[i] * inv; }
      __syncthreads();
      // w = M q
      for (int i = lane; i < n; i += 64) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += M[i + (size_t)n * j] * qv[j];
        wv[i] = s;
      }
      __syncthreads();
      double alpha = 0.0;
      for (int i = lane; i < n; i += 64) alpha += qv[i] * wv[i];
      alpha = wsum(alpha);
      // diag[k] = alpha
      if (lane == k) dreg = alpha;
      for (int i = lane; i < n; i += 64) rv[i] = (wv[i] - qv[i] * alpha) - qold[i] * beta;
      if (kk >= sqrt((double)k) || k == n - 1 || k > sqrt((double)n + 9)) {
        kk = 0;
        double d = dreg, e = ereg;
        if (lane > k) { d = 0.0; e = 0.0; }
        __syncthreads();
        int info = ql_wave(k + 1, d, e, Zs);
        __syncthreads();
        if (info > 0) break;
        mn_old = mn;
        mn = __shfl(d, 0, 64);
        error = Zs[k + 65 * 0];
      }
      __syncthreads();
      double v2 = 0.0;
      for (int i = lane; i < n; i += 64) v2 += rv[i] * rv[i];
      beta = sqrt(wsum(v2));
      if (lane == k) ereg = beta;
      ++k;
      ++kk;
    }
    result = mn - fabs(error * beta);
  }
  if (lane == 0) out[which * nblk + q] = result;
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
    if (int e = c->dw.ensure_items(4 * nb)) return e;
    SDPA_HIP_CHECK(hipMemcpyAsync(c->dw.d_items, it.data(), sizeof(GemmItem) * 4 * nb,
                                  hipMemcpyHostToDevice, c->stream));
    if (int e = gemm_batched(c->stream, c->dw.d_items, 2 * nb, maxn, maxn, 0, 1, 1.0, 0.0, 0))
      return e;
    if (int e = gemm_batched(c->stream, c->dw.d_items + 2 * nb, 2 * nb, maxn, maxn, 0, 0, 1.0, 0.0, 0))
      return e;
    hipLaunchKernelGGL(k_min_eig, dim3(nb, 2), dim3(64), 0, c->stream, c->d_w1, c->lpOff,
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
