// dense.hip -- FP64 dense kernels for CDNA4 (gfx950): a batched MFMA GEMM,
// the blocked right-looking Cholesky used for the Schur matrix (with SDPA's
// adjust rule, sdpa_dpotrf.cpp:31-33) and for the X/Z blocks (LAPACK
// semantics, sdpa_linear.cpp:465-515), the triangular inverse of the X/Z
// factors (sdpa_linear.cpp:566-589) and the two triangular solves of the
// Schur system (sdpa_linear.cpp:843-862).
//
// Tiles are 64x64 with 256 threads = 4 wave64s; each wave owns a 32x32
// quarter computed as 2x2 v_mfma_f64_16x16x4f64 tiles.  The MFMA's A operand
// is fed from B and its B operand from A, so the 16x16 result is C^T and
// lane%16 walks down a column of the column-major C (coalesced stores).
#include "common.h"

#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>

namespace sdpa_amd {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int TB = 64;    // tile edge
constexpr int TLD = 65;   // LDS leading dimension (pad)
constexpr int BK = 16;    // k-depth per LDS stage in the streaming GEMM

// ---------------------------------------------------------------------------
// acc += A(64 x kdepth) * B(kdepth x 64) from LDS tiles stored column-major
// with leading dimension lda_s / ldb_s.  Wave w owns rows [32*(w&1), +32) and
// columns [32*(w>>1), +32) of the 64x64 output; acc[ti][tj] is the 16x16 MFMA
// tile at row block ti, column block tj of that quarter.
__device__ __forceinline__ void mma_tile(v4d acc[2][2], const double* As, int lda_s,
                                         const double* Bs, int ldb_s, int kdepth) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int r0 = 32 * (w & 1), c0 = 32 * (w >> 1);
  const int li = lane & 15, lk = lane >> 4;
  for (int kk = 0; kk < kdepth; kk += 4) {
    double a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      // MFMA A operand <- B[k][col]   (C column = lane%16)
      a[t] = Bs[(kk + lk) + ldb_s * (c0 + 16 * t + li)];
      // MFMA B operand <- A[row][k]   (C row = lane%16)
      b[t] = As[(r0 + 16 * t + li) + lda_s * (kk + lk)];
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
        acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[tj], b[ti], acc[ti][tj], 0, 0, 0);
  }
}

// Position of accumulator element v of tile (ti,tj) for this lane.  The
// f64 16x16x4 MFMA leaves D[i][j] in lane j + 16*(i%4), register i/4
// (measured on gfx950), and D = C^T here, so
// row = r0 + 16*ti + lane%16, col = c0 + 16*tj + 4*v + lane/16.
__device__ __forceinline__ void acc_pos(int ti, int tj, int v, int& r, int& c) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  r = 32 * (w & 1) + 16 * ti + (lane & 15);
  c = 32 * (w >> 1) + 16 * tj + 4 * v + (lane >> 4);
}

__device__ __forceinline__ void acc_zero(v4d acc[2][2]) {
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
}

// ---------------------------------------------------------------------------
// Batched GEMM.  grid = (ceil(maxM/64), ceil(maxN/64), nitems).
template <int TA, int TBt>
__global__ __launch_bounds__(256) void k_gemm(const GemmItem* __restrict__ items, double alpha,
                                              double beta, int lower_only) {
  const GemmItem it = items[blockIdx.z];
  const int bm = blockIdx.x * TB, bn = blockIdx.y * TB;
  if (bm >= it.M || bn >= it.N) return;
  if (lower_only && bm + TB - 1 < bn) return;
  __shared__ double As[TB * (BK + 1)];  // 64 rows x 16 k, ld = 64 (row fastest)
  __shared__ double Bs[BK * TLD];       // 16 k x 64 cols, ld = 16... stored k fastest
  v4d acc[2][2];
  acc_zero(acc);
  const int tid = threadIdx.x;
  for (int k0 = 0; k0 < it.K; k0 += BK) {
    // A tile: op(A)[bm + r][k0 + k], r < 64, k < 16  -> As[r + 64*k]
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int idx = tid + 256 * e;  // 0..1023
      int r, k;
      if (TA == 0) { r = idx & 63; k = idx >> 6; }
      else { k = idx & 15; r = idx >> 4; }
      int gr = bm + r, gk = k0 + k;
      double v = 0.0;
      if (gr < it.M && gk < it.K)
        v = TA == 0 ? it.A[gr + (size_t)it.lda * gk] : it.A[gk + (size_t)it.lda * gr];
      As[r + 64 * k] = v;
    }
    // B tile: op(B)[k0 + k][bn + c] -> Bs[k + 16*c]
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int idx = tid + 256 * e;
      int k, c;
      if (TBt == 0) { k = idx & 15; c = idx >> 4; }
      else { c = idx & 63; k = idx >> 6; }
      int gk = k0 + k, gc = bn + c;
      double v = 0.0;
      if (gk < it.K && gc < it.N)
        v = TBt == 0 ? it.B[gk + (size_t)it.ldb * gc] : it.B[gc + (size_t)it.ldb * gk];
      Bs[k + 16 * c] = v;
    }
    __syncthreads();
    mma_tile(acc, As, 64, Bs, 16, BK);
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        acc_pos(ti, tj, v, r, c);
        int gr = bm + r, gc = bn + c;
        if (gr < it.M && gc < it.N && (!lower_only || gr >= gc)) {
          double* p = it.C + gr + (size_t)it.ldc * gc;
          double val = alpha * acc[ti][tj][v];
          *p = (beta == 0.0) ? val : val + beta * (*p);
        }
      }
}

int gemm_batched(hipStream_t s, const GemmItem* d_items, int nitems, int maxM, int maxN,
                 int transA, int transB, double alpha, double beta, int lower_only) {
  if (nitems <= 0 || maxM <= 0 || maxN <= 0) return 0;
  dim3 grid((maxM + TB - 1) / TB, (maxN + TB - 1) / TB, nitems);
  if (transA == 0 && transB == 0)
    hipLaunchKernelGGL((k_gemm<0, 0>), grid, dim3(256), 0, s, d_items, alpha, beta, lower_only);
  else if (transA == 0 && transB == 1)
    hipLaunchKernelGGL((k_gemm<0, 1>), grid, dim3(256), 0, s, d_items, alpha, beta, lower_only);
  else if (transA == 1 && transB == 0)
    hipLaunchKernelGGL((k_gemm<1, 0>), grid, dim3(256), 0, s, d_items, alpha, beta, lower_only);
  else
    hipLaunchKernelGGL((k_gemm<1, 1>), grid, dim3(256), 0, s, d_items, alpha, beta, lower_only);
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

int DenseWork::ensure_items(int n) {
  if (n <= cap_items) return 0;
  if (d_items) SDPA_HIP_CHECK(hipFree(d_items));
  cap_items = n < 64 ? 64 : n;
  SDPA_HIP_CHECK(hipMalloc(&d_items, sizeof(GemmItem) * cap_items));
  return 0;
}
int DenseWork::ensure_mats(int n) {
  if (n <= cap_mats) return 0;
  if (d_mats) SDPA_HIP_CHECK(hipFree(d_mats));
  cap_mats = n < 64 ? 64 : n;
  SDPA_HIP_CHECK(hipMalloc(&d_mats, sizeof(MatDesc) * cap_mats));
  return 0;
}
int DenseWork::ensure_dinv(size_t n) {
  if (n <= cap_dinv) return 0;
  if (d_dinv) SDPA_HIP_CHECK(hipFree(d_dinv));
  cap_dinv = n;
  SDPA_HIP_CHECK(hipMalloc(&d_dinv, sizeof(double) * cap_dinv));
  return 0;
}
int DenseWork::ensure_vec(size_t n) {
  if (n <= cap_vec) return 0;
  if (d_vec) SDPA_HIP_CHECK(hipFree(d_vec));
  cap_vec = n;
  SDPA_HIP_CHECK(hipMalloc(&d_vec, sizeof(double) * cap_vec));
  return 0;
}
void DenseWork::release() {
  if (d_vec) (void)hipFree(d_vec);
  d_vec = nullptr; cap_vec = 0;
  if (d_items) (void)hipFree(d_items);
  if (d_mats) (void)hipFree(d_mats);
  if (d_dinv) (void)hipFree(d_dinv);
  d_items = nullptr; d_mats = nullptr; d_dinv = nullptr;
  cap_items = cap_mats = 0; cap_dinv = 0;
}

int gemm(hipStream_t s, DenseWork* w, const double* A, int lda, const double* B, int ldb,
         double* C, int ldc, int M, int N, int K, int transA, int transB, double alpha,
         double beta, int lower_only) {
  if (M <= 0 || N <= 0) return 0;
  if (int e = w->ensure_items(1)) return e;
  GemmItem it{A, B, C, M, N, K, lda, ldb, ldc};
  SDPA_HIP_CHECK(hipMemcpyAsync(w->d_items, &it, sizeof(it), hipMemcpyHostToDevice, s));
  return gemm_batched(s, w->d_items, 1, M, N, transA, transB, alpha, beta, lower_only);
}

// ---------------------------------------------------------------------------
// Cholesky of the kb x kb diagonal block at (k0, k0) of every matrix in the
// batch, and the inverse of that factor into dinv (64x64, zero padded) at
// slot (item, k0/64).  One 256-thread work-group per item, the block in LDS.
// The 64 columns go in four 16-wide mini-panels: wave 0 factors a mini-panel
// with the rows in registers and cross-lane broadcasts by readlane (no LDS
// round trip per column); then all four waves apply the mini-panel to the
// trailing block.  The inverse is built the same way: the four 16x16
// diagonal inverses by forward substitution, then block rows
// W_ij = -W_ii sum_k L_ik W_kj.  Pivot rules: mode 0 = dpotrf (d > 0 else
// fail), mode 1 = rATL_dpotrfL (d < -1e-6 fails, d < 1e-14 becomes 1e100).
// Rows/columns beyond kb are padded with the identity.
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_potrf_diag(const MatDesc* __restrict__ mats, int k0,
                                                    int* __restrict__ info,
                                                    double* __restrict__ dinv, int nslot) {
  const int b = blockIdx.x;
  const MatDesc M = mats[b];
  if (M.n <= k0) return;
  if (info[b] != 0) return;
  const int kb = min(TB, M.n - k0);
  __shared__ double T[TB * TLD];
  __shared__ double W[TB * TLD];
  __shared__ double S[16 * 17 * 3];
  __shared__ int failc_s;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double* A = M.ptr + k0 + (size_t)M.ld * k0;
  for (int e = t; e < TB * TB; e += 256) {
    const int r = e & 63, c = e >> 6;
    T[r + TLD * c] = (r < kb && c < kb) ? A[r + (size_t)M.ld * c] : (r == c ? 1.0 : 0.0);
  }
  if (t == 0) failc_s = 0;
  __syncthreads();
  for (int p = 0; p < 4; ++p) {
    const int c0 = 16 * p;
    if (wv == 0) {
      const int r = lane;
      double v[16];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) v[jj] = T[r + TLD * (c0 + jj)];
      int failc = failc_s;
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        const int c = c0 + jj;
        double d = readlane_d(v[jj], c);
        const bool bad = (MODE == 0) ? !(d > 0.0) : (d < -1.0e-6);
        if (MODE == 1) d = (d < 1.0e-14) ? 1.0e+100 : d;
        failc = (failc == 0 && c < kb && bad) ? c + 1 : failc;
        d = failc ? 1.0 : d;
        const double l = sqrt(d);
        const double lrc = (r == c) ? l : v[jj] / l;
        v[jj] = lrc;
#pragma unroll
        for (int kk = jj + 1; kk < 16; ++kk) v[kk] -= lrc * readlane_d(lrc, c0 + kk);
      }
#pragma unroll
      for (int jj = 0; jj < 16; ++jj)
        if (r >= c0 + jj) T[r + TLD * (c0 + jj)] = v[jj];
      if (r == 0) failc_s = failc;
    }
    __syncthreads();
    // trailing update of rows/columns >= c0 + 16 (lower part)
    const int o = c0 + 16, nt = TB - o;
    const int npairs = nt * (nt + 1) / 2;
    for (int e = t; e < npairs; e += 256) {
      // decode e -> (i, j) with i >= j in the trailing block (row-major lower)
      int i = (int)((sqrtf(8.0f * e + 1.0f) - 1.0f) * 0.5f);
      while ((i + 1) * (i + 2) / 2 <= e) ++i;
      while (i * (i + 1) / 2 > e) --i;
      const int j = e - i * (i + 1) / 2;
      const int gi = o + i, gj = o + j;
      double acc = 0.0;
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) acc += T[gi + TLD * (c0 + kk)] * T[gj + TLD * (c0 + kk)];
      T[gi + TLD * gj] -= acc;
    }
    __syncthreads();
  }
  if (failc_s) {
    if (t == 0) info[b] = k0 + failc_s;
    return;
  }
  // write the factor back (lower part of the block)
  for (int e = t; e < TB * TB; e += 256) {
    const int r = e & 63, c = e >> 6;
    if (r >= c && r < kb && c < kb) A[r + (size_t)M.ld * c] = T[r + TLD * c];
  }
  // inverse: four 16x16 diagonal inverses (thread = block bi, column jj)
  for (int e = t; e < TB * TB; e += 256) W[(e & 63) + TLD * (e >> 6)] = 0.0;
  __syncthreads();
  if (t < 64) {
    const int bi = t >> 4, jj = t & 15, o = 16 * bi;
    double w[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      double acc = (i == jj) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) acc -= T[(o + i) + TLD * (o + k)] * w[k];
      w[i] = (i >= jj) ? acc / T[(o + i) + TLD * (o + i)] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) W[(o + i) + TLD * (o + jj)] = w[i];
  }
  __syncthreads();
  // block rows: W_ij = -W_ii * (sum_{k=j}^{i-1} L_ik W_kj), j < i
  for (int bi = 1; bi < 4; ++bi) {
    // S_j = sum_k L_ik W_kj for all j < bi  (bi blocks of 16x16, <= 768 entries)
    for (int e = t; e < bi * 256; e += 256) {
      const int bj = e >> 8, a = (e >> 4) & 15, cc = e & 15;
      double acc = 0.0;
      for (int k = 16 * bj; k < 16 * bi; ++k)
        acc += T[(16 * bi + a) + TLD * k] * W[k + TLD * (16 * bj + cc)];
      S[bj * 272 + a * 17 + cc] = acc;
    }
    __syncthreads();
    for (int e = t; e < bi * 256; e += 256) {
      const int bj = e >> 8, a = (e >> 4) & 15, cc = e & 15;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k) acc += W[(16 * bi + a) + TLD * (16 * bi + k)] * S[bj * 272 + k * 17 + cc];
      W[(16 * bi + a) + TLD * (16 * bj + cc)] = -acc;
    }
    __syncthreads();
  }
  double* Dv = dinv + ((size_t)b * nslot + k0 / TB) * TB * TB;
  for (int e = t; e < TB * TB; e += 256) Dv[e] = W[(e & 63) + TLD * (e >> 6)];
}

// Rows below the diagonal block: L21 := A21 * inv(L11)^T, in place, 64 rows
// per work-group.  grid = (ceil(maxRows/64), nitems).
__global__ __launch_bounds__(256) void k_trsm_panel(const MatDesc* __restrict__ mats, int k0,
                                                    const int* __restrict__ info,
                                                    const double* __restrict__ dinv, int nslot) {
  const int b = blockIdx.y;
  const MatDesc M = mats[b];
  if (M.n <= k0 + TB) return;
  if (info[b] != 0) return;
  const int kb = TB;  // a full 64-wide panel exists whenever rows remain below it
  const int rows = M.n - k0 - kb;
  const int rb = blockIdx.x * TB;
  if (rb >= rows) return;
  __shared__ double As[TB * TLD];
  __shared__ double Bs[TB * TLD];
  double* A21 = M.ptr + (k0 + kb + rb) + (size_t)M.ld * k0;
  const double* Dv = dinv + ((size_t)b * nslot + k0 / TB) * TB * TB;
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    int r = e & 63, c = e >> 6;
    As[r + TLD * c] = (rb + r < rows) ? A21[r + (size_t)M.ld * c] : 0.0;
    // B = inv(L11)^T : B[k][c] = Dv[c][k]
    Bs[c + TLD * r] = Dv[r + TB * c];
  }
  __syncthreads();
  v4d acc[2][2];
  acc_zero(acc);
  mma_tile(acc, As, TLD, Bs, TLD, TB);
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        acc_pos(ti, tj, v, r, c);
        if (rb + r < rows) A21[r + (size_t)M.ld * c] = acc[ti][tj][v];
      }
}

// Update C[r, c] -= sum_{k < kw} L[r, k0 + k] L[c, k0 + k] for rows r >= c
// (both >= off = k0 + kw, r < n) and columns c < colEnd: the trailing update
// of an outer panel (kw = 256, colEnd = n) or the update of the rest of the
// panel after an inner 64-column step (kw = 64, colEnd = panel end).
// 64 x 64 output tile per work-group, k in LDS chunks of 64.
// grid = (row tiles, column tiles, nitems).
__global__ __launch_bounds__(256) void k_syrk_update(const MatDesc* __restrict__ mats, int k0,
                                                     int kw, int colEnd,
                                                     const int* __restrict__ info) {
  const int b = blockIdx.z;
  const MatDesc M = mats[b];
  if (info[b] != 0) return;
  const int off = k0 + kw;
  const int rows = M.n - off;
  const int cend = min(colEnd, M.n) - off;
  const int bm = blockIdx.x * TB, bn = blockIdx.y * TB;
  if (rows <= 0 || bm >= rows || bn >= cend || bm + TB - 1 < bn) return;
  __shared__ double As[TB * TLD];
  __shared__ double Bs[TB * TLD];
  const double* P = M.ptr + off + (size_t)M.ld * k0;  // L[off.., k0..]
  v4d acc[2][2];
  acc_zero(acc);
  for (int kc = 0; kc < kw; kc += TB) {
    for (int e = threadIdx.x; e < TB * TB; e += 256) {
      const int r = e & 63, k = e >> 6;
      const double* col = P + (size_t)M.ld * (kc + k);
      As[r + TLD * k] = (bm + r < rows) ? col[bm + r] : 0.0;
      Bs[k + TLD * r] = (bn + r < cend) ? col[bn + r] : 0.0;
    }
    __syncthreads();
    mma_tile(acc, As, TLD, Bs, TLD, TB);
    __syncthreads();
  }
  double* C = M.ptr + off + (size_t)M.ld * off;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        acc_pos(ti, tj, v, r, c);
        const int gr = bm + r, gc = bn + c;
        if (gr < rows && gc < cend && gr >= gc) C[gr + (size_t)M.ld * gc] -= acc[ti][tj][v];
      }
}

int potrf_batched(hipStream_t s, DenseWork* w, const MatDesc* h_items, int nitems, int mode,
                  int* d_info, double* dinv_out) {
  if (nitems <= 0) return 0;
  int maxn = 0;
  for (int b = 0; b < nitems; ++b) maxn = h_items[b].n > maxn ? h_items[b].n : maxn;
  const int nslot = (maxn + TB - 1) / TB;
  if (int e = w->ensure_mats(nitems)) return e;
  double* dinv = dinv_out;
  if (!dinv) {
    if (int e = w->ensure_dinv((size_t)nitems * nslot * TB * TB)) return e;
    dinv = w->d_dinv;
  }
  SDPA_HIP_CHECK(hipMemcpyAsync(w->d_mats, h_items, sizeof(MatDesc) * nitems,
                                hipMemcpyHostToDevice, s));
  SDPA_HIP_CHECK(hipMemsetAsync(d_info, 0, sizeof(int) * nitems, s));
  // outer panels of PW = 256 columns; inside, 64-column steps whose updates
  // stay within the panel; then one trailing update with K = 256.
  constexpr int PW = 256;
  for (int K0 = 0; K0 < maxn; K0 += PW) {
    const int pend = std::min(K0 + PW, maxn);
    for (int k0 = K0; k0 < pend; k0 += TB) {
      if (mode == 0)
        hipLaunchKernelGGL(k_potrf_diag<0>, dim3(nitems), dim3(256), 0, s, w->d_mats, k0, d_info,
                           dinv, nslot);
      else
        hipLaunchKernelGGL(k_potrf_diag<1>, dim3(nitems), dim3(256), 0, s, w->d_mats, k0, d_info,
                           dinv, nslot);
      const int rows = maxn - k0 - TB;
      if (rows <= 0) break;
      const int rtiles = (rows + TB - 1) / TB;
      hipLaunchKernelGGL(k_trsm_panel, dim3(rtiles, nitems), dim3(256), 0, s, w->d_mats, k0,
                         d_info, dinv, nslot);
      const int pcols = pend - (k0 + TB);
      if (pcols > 0)
        hipLaunchKernelGGL(k_syrk_update, dim3(rtiles, (pcols + TB - 1) / TB, nitems), dim3(256), 0,
                           s, w->d_mats, k0, TB, K0 + PW, d_info);
    }
    const int trail = maxn - (K0 + PW);
    if (trail > 0) {
      const int tt = (trail + TB - 1) / TB;
      hipLaunchKernelGGL(k_syrk_update, dim3(tt, tt, nitems), dim3(256), 0, s, w->d_mats, K0, PW,
                         maxn, d_info);
    }
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Triangular inverse W = L^-1 by block rows: W[I,I] = inv(L_II) and, for
// J < I, W[I,J] = -inv(L_II) * sum_{K=J}^{I-1} L[I,K] W[K,J].
// Step 1: inverses of all diagonal blocks (one wave per (item, block)).
__global__ __launch_bounds__(64) void k_diag_inv(const MatDesc* __restrict__ L,
                                                 const MatDesc* __restrict__ W, int nblk_max,
                                                 double* __restrict__ dinv) {
  const int b = blockIdx.y, I = blockIdx.x;
  const MatDesc Lm = L[b];
  const int k0 = I * TB;
  if (k0 >= Lm.n) return;
  const int kb = min(TB, Lm.n - k0);
  __shared__ double T[TB * TLD];
  __shared__ double V[TB * TLD];
  const int r = threadIdx.x;
  const double* A = Lm.ptr + k0 + (size_t)Lm.ld * k0;
  for (int c = 0; c < kb; ++c) T[r + TLD * c] = (r < kb) ? A[r + (size_t)Lm.ld * c] : 0.0;
  __syncthreads();
  const int j = r;
  for (int i = 0; i < TB; ++i) V[i + TLD * j] = 0.0;
  if (j < kb) {
    for (int i = j; i < kb; ++i) {
      double acc = (i == j) ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) acc -= T[i + TLD * k] * V[k + TLD * j];
      V[i + TLD * j] = acc / T[i + TLD * i];
    }
  }
  __syncthreads();
  double* Dv = dinv + ((size_t)b * nblk_max + I) * TB * TB;
  for (int c = 0; c < TB; ++c) Dv[r + TB * c] = V[r + TLD * c];
  // diagonal block of W and zero the strictly upper blocks of this block row
  const MatDesc Wm = W[b];
  double* Wd = Wm.ptr + k0 + (size_t)Wm.ld * k0;
  for (int c = 0; c < kb; ++c)
    if (r < kb) Wd[r + (size_t)Wm.ld * c] = V[r + TLD * c];
  for (int c = k0 + kb; c < Wm.n; ++c)
    if (r < kb) Wm.ptr[k0 + r + (size_t)Wm.ld * c] = 0.0;
}

// Step 2 (block row I): W[I,J] for every J < I.  grid = (I, nitems).
__global__ __launch_bounds__(256) void k_trtri_row(const MatDesc* __restrict__ L,
                                                   const MatDesc* __restrict__ W, int I,
                                                   int nblk_max, const double* __restrict__ dinv) {
  const int b = blockIdx.y, J = blockIdx.x;
  const MatDesc Lm = L[b], Wm = W[b];
  const int i0 = I * TB;
  if (i0 >= Lm.n) return;
  const int ib = min(TB, Lm.n - i0);
  __shared__ double As[TB * TLD];
  __shared__ double Bs[TB * TLD];
  v4d acc[2][2];
  acc_zero(acc);
  for (int K = J; K < I; ++K) {
    const int k0 = K * TB;
    for (int e = threadIdx.x; e < TB * TB; e += 256) {
      int r = e & 63, c = e >> 6;
      As[r + TLD * c] = (r < ib) ? Lm.ptr[i0 + r + (size_t)Lm.ld * (k0 + c)] : 0.0;
      Bs[r + TLD * c] = Wm.ptr[k0 + r + (size_t)Wm.ld * (J * TB + c)];
    }
    __syncthreads();
    mma_tile(acc, As, TLD, Bs, TLD, TB);
    __syncthreads();
  }
  // S -> LDS (As), then W[I,J] = -inv(L_II) * S
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        acc_pos(ti, tj, v, r, c);
        Bs[r + TLD * c] = acc[ti][tj][v];
      }
  const double* Dv = dinv + ((size_t)b * nblk_max + I) * TB * TB;
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    int r = e & 63, c = e >> 6;
    As[r + TLD * c] = Dv[r + TB * c];
  }
  __syncthreads();
  acc_zero(acc);
  mma_tile(acc, As, TLD, Bs, TLD, TB);
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        acc_pos(ti, tj, v, r, c);
        if (r < ib) Wm.ptr[i0 + r + (size_t)Wm.ld * (J * TB + c)] = -acc[ti][tj][v];
      }
}

int trtri_batched(hipStream_t s, DenseWork* w, const MatDesc* h_L, const MatDesc* h_W,
                  int nitems) {
  if (nitems <= 0) return 0;
  int maxn = 0;
  for (int b = 0; b < nitems; ++b) maxn = h_L[b].n > maxn ? h_L[b].n : maxn;
  const int nblk = (maxn + TB - 1) / TB;
  if (int e = w->ensure_mats(2 * nitems)) return e;
  if (int e = w->ensure_dinv((size_t)nitems * nblk * TB * TB)) return e;
  SDPA_HIP_CHECK(hipMemcpyAsync(w->d_mats, h_L, sizeof(MatDesc) * nitems, hipMemcpyHostToDevice, s));
  SDPA_HIP_CHECK(hipMemcpyAsync(w->d_mats + nitems, h_W, sizeof(MatDesc) * nitems,
                                hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_diag_inv, dim3(nblk, nitems), dim3(64), 0, s, w->d_mats,
                     w->d_mats + nitems, nblk, w->d_dinv);
  for (int I = 1; I < nblk; ++I)
    hipLaunchKernelGGL(k_trtri_row, dim3(I, nitems), dim3(256), 0, s, w->d_mats,
                       w->d_mats + nitems, I, nblk, w->d_dinv);
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Solves with the m x m lower factor L (forward L t = x, backward L^T x = t)
// by steps of SB = 256 rows, with the inverses W_b of the 256 x 256 diagonal
// blocks precomputed (diag_block_inverses, once per factorization).  Step b
// is one launch: work-group 0 applies block column b to the rows of block b+1
// and then forms x_{b+1} = W_{b+1} y_{b+1}; the other work-groups apply block
// column b to the rows beyond.  A step only reads x_b (written by the
// previous launch) and never reads rows another work-group of the same launch
// writes.
constexpr int SB = 256;

int diag_block_inverses(hipStream_t s, DenseWork* w, const double* L, int ld, int m,
                        double* Winv) {
  const int nb = (m + SB - 1) / SB;
  std::vector<MatDesc> Ls(nb), Ws(nb);
  for (int b = 0; b < nb; ++b) {
    const int k0 = b * SB;
    Ls[b] = MatDesc{const_cast<double*>(L) + k0 + (size_t)ld * k0, std::min(SB, m - k0), ld};
    Ws[b] = MatDesc{Winv + (size_t)b * SB * SB, std::min(SB, m - k0), SB};
  }
  return trtri_batched(s, w, Ls.data(), Ws.data(), nb);
}

// Solve step, part 1: x_b = W_b y_b (trans 0) or W_b^T y_b (trans 1), by
// 8 work-groups of 32 outputs each; each thread keeps 8 independent loads in
// flight (the W block streams from HBM).  grid = 8.
__global__ __launch_bounds__(256) void k_block_gemv(const double* __restrict__ Winv, int m, int b,
                                                    int trans, const double* __restrict__ y,
                                                    double* __restrict__ sol) {
  __shared__ double v[SB];
  __shared__ double part[256];
  const int k0 = b * SB, kb = min(SB, m - k0);
  const int t = threadIdx.x;
  v[t] = (t < kb) ? y[k0 + t] : 0.0;
  __syncthreads();
  const double* W = Winv + (size_t)b * SB * SB;
  const int i0 = blockIdx.x * 32;
  double acc = 0.0;
  if (!trans) {
    // thread = (output i0 + (t & 31), column chunk (t >> 5) of 32 columns)
    const int i = i0 + (t & 31), cb = (t >> 5) * 32;
    if (i < kb) {
      double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int c = 0; c < 32; ++c)
        if (cb + c < kb) a[c & 7] += W[i + (size_t)SB * (cb + c)] * v[cb + c];
      acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
    part[t] = acc;
    __syncthreads();
    if (t < 32 && i0 + t < kb) {
      double s2 = 0.0;
      for (int q = 0; q < 8; ++q) s2 += part[q * 32 + t];
      sol[k0 + i0 + t] = s2;
    }
  } else {
    // thread = (output i0 + (t >> 3), k chunk (t & 7) of 32): column i of W is contiguous
    const int i = i0 + (t >> 3), kc = (t & 7) * 32;
    if (i < kb) {
      double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int k = 0; k < 32; ++k)
        if (kc + k < kb) a[k & 7] += W[(kc + k) + (size_t)SB * i] * v[kc + k];
      acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
    // reduce the 8 chunks of each output (consecutive lanes)
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    acc += __shfl_xor(acc, 4, 64);
    if ((t & 7) == 0 && i < kb) sol[k0 + i] = acc;
  }
}

// Solve step, part 2 (forward): y[r] -= L[r, block b] x_b for the rows after
// block b, 64 rows x 4 column quarters per work-group.
__global__ __launch_bounds__(256) void k_fwd_update(const double* __restrict__ L, int ld, int m,
                                                    int b, const double* __restrict__ sol,
                                                    double* __restrict__ y) {
  __shared__ double xb[SB];
  __shared__ double part[256];
  const int k0 = b * SB, kb = min(SB, m - k0);
  const int t = threadIdx.x, i = t & 63, q = t >> 6;
  xb[t] = (t < kb) ? sol[k0 + t] : 0.0;
  __syncthreads();
  const int r = k0 + kb + blockIdx.x * 64 + i;
  double acc = 0.0;
  if (r < m) {
    const double* Lr = L + r + (size_t)ld * (k0 + q * 64);
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int ce = min(64, kb - q * 64);
#pragma unroll
    for (int c = 0; c < 64; ++c)
      if (c < ce) a[c & 7] += Lr[(size_t)ld * c] * xb[q * 64 + c];
    acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  }
  part[t] = acc;
  __syncthreads();
  if (q == 0 && r < m) y[r] -= (part[i] + part[64 + i]) + (part[128 + i] + part[192 + i]);
}

// Solve step, part 2 (backward): y[r] -= sum_c L[k0 + c, r] x_b[c] for the
// rows before block b; one wave per row (column r of L below k0 is contiguous).
__global__ __launch_bounds__(256) void k_bwd_update(const double* __restrict__ L, int ld, int m,
                                                    int b, const double* __restrict__ sol,
                                                    double* __restrict__ y) {
  __shared__ double xb[SB];
  const int k0 = b * SB, kb = min(SB, m - k0);
  const int t = threadIdx.x, lane = t & 63;
  xb[t] = (t < kb) ? sol[k0 + t] : 0.0;
  __syncthreads();
  const int r = blockIdx.x * 4 + (t >> 6);
  if (r >= k0) return;
  const double* Lc = L + k0 + (size_t)ld * r;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (lane < kb) a0 = Lc[lane] * xb[lane];
  if (lane + 64 < kb) a1 = Lc[lane + 64] * xb[lane + 64];
  if (lane + 128 < kb) a2 = Lc[lane + 128] * xb[lane + 128];
  if (lane + 192 < kb) a3 = Lc[lane + 192] * xb[lane + 192];
  double acc = (a0 + a1) + (a2 + a3);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) y[r] -= acc;
}

int chol_solve(hipStream_t s, DenseWork* w, const double* L, int ld, int m, const double* Winv,
               double* x) {
  if (m <= 0) return 0;
  if (int e = w->ensure_vec((size_t)m)) return e;
  double* t = w->d_vec;
  const int nb = (m + SB - 1) / SB;
  // forward: t = L^-1 x   (x is consumed as the right-hand side)
  for (int b = 0; b < nb; ++b) {
    hipLaunchKernelGGL(k_block_gemv, dim3(8), dim3(256), 0, s, Winv, m, b, 0, x, t);
    const int beyond = m - (b + 1) * SB;
    if (beyond > 0)
      hipLaunchKernelGGL(k_fwd_update, dim3((beyond + 63) / 64), dim3(256), 0, s, L, ld, m, b, t, x);
  }
  // backward: x = L^-T t   (t is consumed)
  for (int b = nb - 1; b >= 0; --b) {
    hipLaunchKernelGGL(k_block_gemv, dim3(8), dim3(256), 0, s, Winv, m, b, 1, t, x);
    if (b > 0)
      hipLaunchKernelGGL(k_bwd_update, dim3((b * SB + 3) / 4), dim3(256), 0, s, L, ld, m, b, x, t);
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace sdpa_amd
