// eig.hip -- step-length eigenvalue estimates on the device
// (StepLength::computeStepLength -> Jal::getMinEigen, sdpa_parts.cpp:363-432,
// sdpa_jordan.cpp:58-109).
//
// For each SDP block the scaled direction M = L D L^T (L = inverse Cholesky
// factor, D = Dx or Dz) is formed by two batched MFMA GEMMs (or library
// GEMMs for one large block).  Then, per (block, X/Z):
//   n > 20 : one 256-thread work-group runs the truncated Lanczos of
//            Lal::getMinEigen (sdpa_linear.cpp:29-143): start vector of ones,
//            at most min(n, sqrt(n)+10) steps, the tridiagonal eigenproblem
//            re-solved on the schedule of line 84, the stop tests of lines
//            52-56, result min - |err*beta|.  The tridiagonal problem is
//            solved by Sturm multisection with inverse iteration for the last
//            eigenvector component (tridiag_min_block, eig_dev.h);
//   n <= 20: one wave reduces M to tridiagonal form (Householder) and takes
//            its smallest eigenvalue by Sturm multisection (the dsyev path,
//            sdpa_linear.cpp:145-182, which needs only the smallest).
#include "ctx.h"
#include "eig_dev.h"

#include <cmath>

namespace sdpa_amd {

constexpr int kMaxLanczos = 64;   // tridiagonal size limit (sqrt(n)+10 <= 64 for n <= 2916)
constexpr int kMaxN = 2048;       // largest block the single-wave kernel holds in LDS vectors

// n <= 20: one wave per (block, which): Householder reduction of M = L D L^T
// to tridiagonal form and its smallest eigenvalue by Sturm multisection.
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
  __shared__ double dg[64], od[64], trisc[2];
  for (int j = 0; j < n; ++j)
    if (lane < n) Zs[lane + 65 * j] = M[lane + (size_t)n * j];
  __syncthreads();
  // dsyev('L') reads the lower triangle (mirrored inside); its smallest
  // eigenvalue through the tridiagonal form
  householder_tridiag_wave(n, Zs, 65, dg, od);
  __syncthreads();
  double th = 0.0, zl = 0.0;
  tridiag_min_block(n, dg, od, th, zl, 64, trisc, false);
  if (lane == 0) out[which * nblk + q] = th;
}

constexpr int kLzThreads = 256;  // global-memory matrix-vector product (n > kLrMaxN)
constexpr int kLrThreads = 512;  // matrix held in registers (20 < n <= kLrMaxN)
constexpr int kLrCols = 104;     // columns per thread in the register-resident product
constexpr int kLrMaxN = 2 * kLrCols;

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wsum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < NT / 64; ++i) s += red[i];
  return s;
}

// n > 20: the truncated Lanczos of Lal::getMinEigen on M, one work-group of
// NT threads per (block, which).  The matrix-vector product w = M q:
//   NC == 0: one thread per row, streaming M from global memory every step
//            (lanes over consecutive rows: each column read is coalesced);
//   NC > 0 : M stays on chip for the whole run.  Wave w holds rows
//            64 (w % 4) + lane and the column half w / 4 (NC columns, n <= 2
//            NC) in registers, loaded once; a step is NC register FMAs per
//            thread against LDS broadcasts of q, then the two halves are
//            added.  (Per step the streaming form moves all n^2 doubles
//            through one CU's L1: about 8 us at n = 200, against about 1 us.)
// The tridiagonal eigenproblem (tridiag_min_block) takes all NT threads.
template <int NT, int NC>
__global__ __launch_bounds__(NT) void k_lanczos(const double* __restrict__ Ms, long stride,
                                                const long* __restrict__ off,
                                                const int* __restrict__ nn, int nblk,
                                                double* __restrict__ out) {
  constexpr int VN = NC > 0 ? 2 * NC : kMaxN;  // LDS vector length
  const int q = blockIdx.x, which = blockIdx.y;
  const int n = nn[q];
  if (n <= 20 || (NC > 0 && n > 2 * NC) || (NC == 0 && n <= kLrMaxN)) return;
  const double* M = Ms + which * stride + off[q];
  const int tid = threadIdx.x;
  __shared__ double dg[kMaxLanczos], od[kMaxLanczos];  // tridiagonal of the recurrence
  __shared__ double qv[VN], qold[VN], rv[VN], wv[VN];
  __shared__ double part[NC > 0 ? 2 * 256 : 1];
  __shared__ double red[NT / 64];
  __shared__ double sh[4];  // theta, last eigenvector component
  __shared__ double trisc[2 * (NT / 64)];
  // register-resident M (NC > 0): row ri, columns [cj0, cj0 + NC)
  const int wave = tid >> 6, ri = 64 * (wave & 3) + (tid & 63), half = (n + 1) / 2;
  const int cj0 = (wave >> 2) * half, cjn = min(n, cj0 + half) - cj0;
  double mreg[NC > 0 ? NC : 1];
  if (NC > 0) {
    // straight-line loads from clamped indices (cjn >= 10 here), all in
    // flight at once, then zeros outside the block (by a 0/1 factor: a select
    // there becomes a branch around each load)
    const double* Mr = M + min(ri, n - 1) + (size_t)n * cj0;
    const double rowf = ri < n ? 1.0 : 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) mreg[c] = __builtin_nontemporal_load(Mr + (size_t)n * min(c, cjn - 1));
#pragma unroll
    for (int c = 0; c < NC; ++c) mreg[c] *= (c < cjn) ? rowf : 0.0;
  }
  double mn = 1.0e+51, mn_old = 1.0e+52, error = 1.0e+10;
  int k = 0, kk = 0;
  // entries past n stay 0: the register product reads q up to index 2 NC
  for (int i = tid; i < VN; i += NT) { qv[i] = 0.0; rv[i] = i < n ? 1.0 : 0.0; }
  double beta = sqrt((double)n);
  const double sqn = sqrt((double)n);
  while (k < n && k < sqn + 10 && k < kMaxLanczos && beta > 1.0e-16 &&
         (fabs(mn - mn_old) > (1.0e-5) * fabs(mn) + (1.0e-8) ||
          fabs(error * beta) > (1.0e-2) * fabs(mn) + (1.0e-4))) {
    const double inv = 1.0 / beta;
    __syncthreads();
    for (int i = tid; i < n; i += NT) { qold[i] = qv[i]; qv[i] = rv[i] * inv; }
    __syncthreads();
    if (NC > 0) {
      double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < NC; ++c) a[c & 3] += mreg[c] * qv[cj0 + c];
      part[(wave >> 2) * 256 + ri] = (a[0] + a[1]) + (a[2] + a[3]);
      __syncthreads();
      for (int i = tid; i < n; i += NT) wv[i] = part[i] + part[256 + i];
    } else {
      // one thread per row i, lanes over consecutive rows, so each step j
      // reads 64 consecutive entries of column j (coalesced); q_j is an LDS
      // broadcast.  32 columns of loads in flight per thread.
      for (int i = tid; i < n; i += NT) {
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
    }
    __syncthreads();
    double a = 0.0;
    for (int i = tid; i < n; i += NT) a += qv[i] * wv[i];
    const double alpha = block_sum<NT>(a, red);
    if (tid == 0) dg[k] = alpha;
    for (int i = tid; i < n; i += NT) rv[i] = (wv[i] - qv[i] * alpha) - qold[i] * beta;
    if (kk >= sqrt((double)k) || k == n - 1 || k > sqrt((double)n + 9)) {
      kk = 0;
      __syncthreads();
      {
        double th = 0.0, zl = 0.0;
        tridiag_min_block(k + 1, dg, od, th, zl, NT, trisc);
        if (tid == 0) { sh[0] = th; sh[1] = zl; }
      }
      __syncthreads();
      mn_old = mn;
      mn = sh[0];
      error = sh[1];
    }
    double v2 = 0.0;
    for (int i = tid; i < n; i += NT) v2 += rv[i] * rv[i];
    beta = sqrt(block_sum<NT>(v2, red));
    if (tid == 0) od[k] = beta;
    ++k;
    ++kk;
  }
  if (tid == 0) out[which * nblk + q] = mn - fabs(error * beta);
}

// The launches of min_eig and the copies of their results (per SDP block to
// c->h_eig, the LP minima to c->h_red[lp_slot..]); no synchronization.
int min_eig_enqueue(Ctx* c, int lp_slot) {
  const int nb = c->nSDP;
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
        if (int e = dgemm(c->stream, 0, k < 2 ? 1 : 0, g.M, g.N, g.K, 1.0, g.A, g.lda, 0, g.B, g.ldb, 0, 0.0, g.C, g.ldc, 0, 1))
          return e;
      }
    } else {
      const void* di = nullptr;
      if (int e = c->dw.desc_copy(c->stream, it.data(), sizeof(GemmItem) * 4 * nb, &di)) return e;
      const GemmItem* d_items = static_cast<const GemmItem*>(di);
      if (int e = gemm_batched(c->stream, d_items, 2 * nb, maxn, maxn, 0, 1, 1.0, 0.0, 0))
        return e;
      if (int e = gemm_batched(c->stream, d_items + 2 * nb, 2 * nb, maxn, maxn, 0, 0, 1.0,
                               0.0, 0))
        return e;
    }
    hipLaunchKernelGGL(k_min_eig_small, dim3(nb, 2), dim3(64), 0, c->stream, c->d_w1, c->lpOff,
                       c->d_sdpOff, c->d_sdpN, nb, c->d_eig);
    int nlr = 0, nlz = 0;
    for (int q = 0; q < nb; ++q) {
      nlr += c->sdpN[q] > 20 && c->sdpN[q] <= kLrMaxN;
      nlz += c->sdpN[q] > kLrMaxN;
    }
    if (nlr)
      hipLaunchKernelGGL((k_lanczos<kLrThreads, kLrCols>), dim3(nb, 2), dim3(kLrThreads), 0,
                         c->stream, c->d_w1, c->lpOff, c->d_sdpOff, c->d_sdpN, nb, c->d_eig);
    if (nlz)
      hipLaunchKernelGGL((k_lanczos<kLzThreads, 0>), dim3(nb, 2), dim3(kLzThreads), 0, c->stream,
                         c->d_w1, c->lpOff, c->d_sdpOff, c->d_sdpN, nb, c->d_eig);
    SDPA_HIP_CHECK(hipGetLastError());
    SDPA_HIP_CHECK(hipMemcpyAsync(c->h_eig, c->d_eig, sizeof(double) * 2 * nb,
                                  hipMemcpyDeviceToHost, c->stream));
  }
  if (c->nLP > 0)
    if (int e = lp_min_enqueue(c, lp_slot)) return e;
  return 0;
}

// after the stream has passed min_eig_enqueue: the minima over the blocks
void min_eig_finish(const Ctx* c, int lp_slot, double* lamX, double* lamZ) {
  const int nb = c->nSDP;
  double mX = 1.0e50, mZ = 1.0e50;
  for (int q = 0; q < nb; ++q) {
    mX = c->h_eig[q] < mX ? c->h_eig[q] : mX;
    mZ = c->h_eig[nb + q] < mZ ? c->h_eig[nb + q] : mZ;
  }
  if (c->nLP > 0) {
    const double a = c->h_red[lp_slot], b = c->h_red[lp_slot + 1];
    mX = a < mX ? a : mX;
    mZ = b < mZ ? b : mZ;
  }
  *lamX = mX;
  *lamZ = mZ;
}

int min_eig(Ctx* c, double* lamX, double* lamZ) {
  if (int e = min_eig_enqueue(c, kSlotLpMin)) return e;
  SDPA_HIP_CHECK(hipStreamSynchronize(c->stream));
  min_eig_finish(c, kSlotLpMin, lamX, lamZ);
  return 0;
}

}  // namespace sdpa_amd
