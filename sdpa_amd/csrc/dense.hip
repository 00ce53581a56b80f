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
// batch, and the inverse of that factor into dinv (64x64 per item, zero
// padded).  One wave per item.  Pivot rules: mode 0 = dpotrf (d > 0 else
// fail), mode 1 = rATL_dpotrfL (d < -1e-6 fail, d < 1e-14 -> 1e100).
__global__ __launch_bounds__(64) void k_potrf_diag(const MatDesc* __restrict__ mats, int k0,
                                                   int mode, int* __restrict__ info,
                                                   double* __restrict__ dinv) {
  const int b = blockIdx.x;
  const MatDesc M = mats[b];
  if (M.n <= k0) return;
  if (info[b] != 0) return;
  const int kb = min(TB, M.n - k0);
  __shared__ double T[TB * TLD];
  __shared__ double W[TB * TLD];
  __shared__ int fail;
  const int r = threadIdx.x;
  double* A = M.ptr + k0 + (size_t)M.ld * k0;
  for (int c = 0; c < kb; ++c) T[r + TLD * c] = (r < kb) ? A[r + (size_t)M.ld * c] : 0.0;
  if (r == 0) fail = 0;
  __syncthreads();
  for (int c = 0; c < kb; ++c) {
    double s = T[r + TLD * c];
    for (int k = 0; k < c; ++k) s -= T[r + TLD * k] * T[c + TLD * k];
    double d = __shfl(s, c, 64);
    bool bad;
    if (mode == 0) {
      bad = !(d > 0.0);
    } else {
      bad = d < -1.0e-6;
      if (!bad && d < 1.0e-14) d = 1.0e+100;
    }
    if (bad) {
      if (r == 0) { info[b] = k0 + c + 1; fail = 1; }
      break;
    }
    double l = sqrt(d);
    __syncthreads();
    if (r == c) T[c + TLD * c] = l;
    else if (r > c && r < kb) T[r + TLD * c] = s / l;
    __syncthreads();
  }
  __syncthreads();
  if (fail) return;
  // write the factor back (lower part of the block)
  for (int c = 0; c < kb; ++c)
    if (r >= c && r < kb) A[r + (size_t)M.ld * c] = T[r + TLD * c];
  // inverse: lane j owns column j of W = L^-1
  const int j = r;
  for (int i = 0; i < TB; ++i) W[i + TLD * j] = 0.0;
  if (j < kb) {
    for (int i = j; i < kb; ++i) {
      double acc = (i == j) ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) acc -= T[i + TLD * k] * W[k + TLD * j];
      W[i + TLD * j] = acc / T[i + TLD * i];
    }
  }
  __syncthreads();
  double* Dv = dinv + (size_t)b * TB * TB;
  for (int c = 0; c < TB; ++c) Dv[r + TB * c] = W[r + TLD * c];
}

// Rows below the diagonal block: L21 := A21 * inv(L11)^T, in place, 64 rows
// per work-group.  grid = (ceil(maxRows/64), nitems).
__global__ __launch_bounds__(256) void k_trsm_panel(const MatDesc* __restrict__ mats, int k0,
                                                    const int* __restrict__ info,
                                                    const double* __restrict__ dinv) {
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
  const double* Dv = dinv + (size_t)b * TB * TB;
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

// Trailing update A22 -= L21 * L21^T on lower tiles.  grid = (T, T, nitems)
// over the largest trailing size; K = 64.
__global__ __launch_bounds__(256) void k_syrk_update(const MatDesc* __restrict__ mats, int k0,
                                                     const int* __restrict__ info) {
  const int b = blockIdx.z;
  const MatDesc M = mats[b];
  if (M.n <= k0 + TB) return;
  if (info[b] != 0) return;
  const int off = k0 + TB;
  const int rows = M.n - off;
  const int bm = blockIdx.x * TB, bn = blockIdx.y * TB;
  if (bm >= rows || bn >= rows || bm < bn) return;
  __shared__ double As[TB * TLD];
  __shared__ double Bs[TB * TLD];
  const double* L21 = M.ptr + off + (size_t)M.ld * k0;
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    int r = e & 63, k = e >> 6;
    As[r + TLD * k] = (bm + r < rows) ? L21[bm + r + (size_t)M.ld * k] : 0.0;
    // B[k][c] = L21[bn + c][k]
    Bs[k + TLD * r] = (bn + r < rows) ? L21[bn + r + (size_t)M.ld * k] : 0.0;
  }
  __syncthreads();
  v4d acc[2][2];
  acc_zero(acc);
  mma_tile(acc, As, TLD, Bs, TLD, TB);
  double* C = M.ptr + off + (size_t)M.ld * off;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        acc_pos(ti, tj, v, r, c);
        int gr = bm + r, gc = bn + c;
        if (gr < rows && gc < rows && gr >= gc) C[gr + (size_t)M.ld * gc] -= acc[ti][tj][v];
      }
}

int potrf_batched(hipStream_t s, DenseWork* w, const MatDesc* h_items, int nitems, int mode,
                  int* d_info) {
  if (nitems <= 0) return 0;
  if (int e = w->ensure_mats(nitems)) return e;
  if (int e = w->ensure_dinv((size_t)nitems * TB * TB)) return e;
  SDPA_HIP_CHECK(hipMemcpyAsync(w->d_mats, h_items, sizeof(MatDesc) * nitems,
                                hipMemcpyHostToDevice, s));
  SDPA_HIP_CHECK(hipMemsetAsync(d_info, 0, sizeof(int) * nitems, s));
  int maxn = 0;
  for (int b = 0; b < nitems; ++b) maxn = h_items[b].n > maxn ? h_items[b].n : maxn;
  for (int k0 = 0; k0 < maxn; k0 += TB) {
    hipLaunchKernelGGL(k_potrf_diag, dim3(nitems), dim3(64), 0, s, w->d_mats, k0, mode, d_info,
                       w->d_dinv);
    int rows = maxn - k0 - TB;
    if (rows <= 0) break;
    int tiles = (rows + TB - 1) / TB;
    hipLaunchKernelGGL(k_trsm_panel, dim3(tiles, nitems), dim3(256), 0, s, w->d_mats, k0, d_info,
                       w->d_dinv);
    hipLaunchKernelGGL(k_syrk_update, dim3(tiles, tiles, nitems), dim3(256), 0, s, w->d_mats, k0,
                       d_info);
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
// Solves with the m x m lower factor L (in place on x): forward L y = x,
// then backward L^T z = y.  Block size 64; each step is one launch in which
// every work-group first solves the 64x64 diagonal system (redundantly, from
// LDS) and then applies the block column to its share of the remaining rows.
// The right-hand side y is consumed block by block; the solved block goes to
// sol, so no work-group ever reads a value another one writes in the same step.
__global__ __launch_bounds__(256) void k_fwd_step(const double* __restrict__ L, int ld, int m,
                                                  int k0, double* __restrict__ x,
                                                  double* __restrict__ sol) {
  const int kb = min(TB, m - k0);
  __shared__ double T[TB * TLD];
  __shared__ double xs[TB];
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    int r = e & 63, c = e >> 6;
    T[r + TLD * c] = (r < kb && c < kb) ? L[k0 + r + (size_t)ld * (k0 + c)] : 0.0;
  }
  if (threadIdx.x < TB) xs[threadIdx.x] = (threadIdx.x < kb) ? x[k0 + threadIdx.x] : 0.0;
  __syncthreads();
  if (threadIdx.x < 64) {
    const int r = threadIdx.x;
    double xr = xs[r];
    for (int c = 0; c < kb; ++c) {
      double xc = __shfl(xr, c, 64);
      if (r == c) xr = xr / T[c + TLD * c];
      xc = __shfl(xr, c, 64);
      if (r > c) xr -= T[r + TLD * c] * xc;
    }
    xs[r] = xr;
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x < kb) sol[k0 + threadIdx.x] = xs[threadIdx.x];
  const int row = k0 + kb + blockIdx.x * 256 + threadIdx.x;
  if (row < m) {
    double acc = 0.0;
    const double* Lr = L + row + (size_t)ld * k0;
    for (int c = 0; c < kb; ++c) acc += Lr[(size_t)ld * c] * xs[c];
    x[row] -= acc;
  }
}

// Backward step for block k0 (descending): solve L_bb^T x_b = x_b, then
// x_r -= sum_c L[c, r] x_c for rows r < k0 (one wave per row r, lanes over c).
__global__ __launch_bounds__(256) void k_bwd_step(const double* __restrict__ L, int ld, int m,
                                                  int k0, double* __restrict__ x,
                                                  double* __restrict__ sol) {
  const int kb = min(TB, m - k0);
  __shared__ double T[TB * TLD];
  __shared__ double xs[TB];
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    int r = e & 63, c = e >> 6;
    T[r + TLD * c] = (r < kb && c < kb) ? L[k0 + r + (size_t)ld * (k0 + c)] : 0.0;
  }
  if (threadIdx.x < TB) xs[threadIdx.x] = (threadIdx.x < kb) ? x[k0 + threadIdx.x] : 0.0;
  __syncthreads();
  if (threadIdx.x < 64) {
    const int r = threadIdx.x;
    double xr = xs[r];
    for (int c = kb - 1; c >= 0; --c) {
      if (r == c) xr = xr / T[c + TLD * c];
      double xc = __shfl(xr, c, 64);
      if (r < c) xr -= T[c + TLD * r] * xc;  // (L^T)[r][c] = L[c][r]
    }
    xs[r] = xr;
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x < kb) sol[k0 + threadIdx.x] = xs[threadIdx.x];
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row < k0) {
    const double* Lc = L + k0 + (size_t)ld * row;  // L[k0.., row]
    double acc = (lane < kb) ? Lc[lane] * xs[lane] : 0.0;
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) x[row] -= acc;
  }
}

int chol_solve(hipStream_t s, DenseWork* w, const double* L, int ld, int m, double* x) {
  if (m <= 0) return 0;
  if (int e = w->ensure_vec((size_t)m)) return e;
  double* t = w->d_vec;
  for (int k0 = 0; k0 < m; k0 += TB) {  // t = L^-1 x  (x is scratch)
    int rem = m - k0 - TB;
    int g = rem > 0 ? (rem + 255) / 256 : 1;
    hipLaunchKernelGGL(k_fwd_step, dim3(g), dim3(256), 0, s, L, ld, m, k0, x, t);
  }
  const int last = ((m - 1) / TB) * TB;
  for (int k0 = last; k0 >= 0; k0 -= TB) {  // x = L^-T t  (t is scratch)
    int g = k0 > 0 ? (k0 + 3) / 4 : 1;
    hipLaunchKernelGGL(k_bwd_step, dim3(g), dim3(256), 0, s, L, ld, m, k0, t, x);
  }
  SDPA_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace sdpa_amd
